/*
 * crimson_oracle.c -- FP64 CPU restatement of the reference formation path.
 * TEST INFRASTRUCTURE ONLY (see crimson_oracle.h).  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off so every multiply and add rounds separately, as
 * the reference Fortran does without FMA contraction).
 *
 * The element routines keep the reference's block-SoA shape (arrays of npro
 * elements, npro fastest) and evaluate every expression in the reference's
 * left-to-right order, so the arithmetic sequence per element is the one the
 * Fortran source prescribes.
 */
#include "crimson_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* The threaded build (-fopenmp -DCFO_THREADS, liboracle_mt.so) timed as the
 * all-core CPU baseline splits the O(N) loops of memLS, PRECOND and the Val4
 * packing over threads with static schedules, and the formation over element
 * blocks and CSR rows (cfo_elmgmr_interior_plan).  The serial build ignores
 * CFO_PFOR. */
#ifdef CFO_THREADS
#include <omp.h>
#define CFO_PFOR _Pragma("omp parallel for schedule(static)")
#else
#define CFO_PFOR
#endif

/* 4-point rule (symtet.c:22-23,72-74), Qwt = 4/3 * 1/4 (genint.f90:85) */
static const double A4 = 0.5854101966249685;
static const double B4 = 0.1381966011250150;

/* shp(a,q) and shgl(j,a) = dN_a/dxi_j / 2 (uniformP.c:22-30,
 * genshp.f90:35-36) */
static void tet_tables(double shp[4][4], double shgl[4][3], double *qwt) {
  const double qpt[4][3] = {{A4, B4, B4}, {B4, A4, B4}, {B4, B4, A4},
                            {B4, B4, B4}};
  for (int q = 0; q < 4; ++q) {
    shp[0][q] = qpt[q][0];
    shp[1][q] = qpt[q][1];
    shp[2][q] = qpt[q][2];
    shp[3][q] = 1.0e0 - qpt[q][0] - qpt[q][1] - qpt[q][2];
  }
  for (int a = 0; a < 4; ++a)
    for (int j = 0; j < 3; ++j)
      shgl[a][j] = (a == 3) ? -1.0e0 / 2.0 : ((a == j) ? 1.0e0 / 2.0 : 0.0e0 / 2.0);
  *qwt = (4.0 / 3.0) * 0.25;
}

void cfo_time_params(double rhoinf, double *almi, double *alfi, double *gami) {
  /* itrPC.f90:16-37 */
  if (rhoinf < 0.0 || rhoinf > 1.0) {
    *almi = 1.0;
    *alfi = 1.0;
    *gami = 1.0;
  } else {
    *almi = (3.0 - rhoinf) / (1.0 + rhoinf) / 2.0;
    *alfi = 1.0 / (1.0 + rhoinf);
    *gami = 0.5 + *almi - *alfi;
  }
}

/* index helpers for the reference's block arrays (0-based) */
#define YL(e, a, k) yl[(e) + n * ((a) + 4 * (k))]
#define ACL(e, a, k) acl[(e) + n * ((a) + 4 * (k))]
#define XL(e, a, i) xl[(e) + n * ((a) + 4 * (i))]
#define QL(e, a, k) ql[(e) + n * ((a) + 4 * (k))]
#define RL(e, a, k) rl[(e) + n * ((a) + 4 * (k))]
#define KE(e, c, a, b) xKebe[(e) + n * ((c) + 9 * ((a) + 4 * (b)))]
#define GC(e, c, a, b) xGoC[(e) + n * ((c) + 4 * ((a) + 4 * (b)))]

/* e3metric.f90:8-81.  dxdxi accumulates from zero over nodes; the inverse
 * is formed by cofactors; WdetJ = Qwt / tmp; shg = shgl . dxidx. */
static void metric(int n, const double *xl, const double shgl[4][3],
                   double qwt, double (*dxidx)[3][3], double *WdetJ,
                   double (*shg)[4][3]) {
  for (int e = 0; e < n; ++e) {
    double d[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) d[i][j] = 0.0;
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) d[i][j] = d[i][j] + XL(e, a, i) * shgl[a][j];
    double (*x)[3] = dxidx[e];
    x[0][0] = d[1][1] * d[2][2] - d[2][1] * d[1][2];
    x[0][1] = d[2][1] * d[0][2] - d[0][1] * d[2][2];
    x[0][2] = d[0][1] * d[1][2] - d[0][2] * d[1][1];
    double tmp = 1.0 / (x[0][0] * d[0][0] + x[0][1] * d[1][0] + x[0][2] * d[2][0]);
    x[0][0] = x[0][0] * tmp;
    x[0][1] = x[0][1] * tmp;
    x[0][2] = x[0][2] * tmp;
    x[1][0] = (d[1][2] * d[2][0] - d[1][0] * d[2][2]) * tmp;
    x[1][1] = (d[0][0] * d[2][2] - d[2][0] * d[0][2]) * tmp;
    x[1][2] = (d[1][0] * d[0][2] - d[0][0] * d[1][2]) * tmp;
    x[2][0] = (d[1][0] * d[2][1] - d[1][1] * d[2][0]) * tmp;
    x[2][1] = (d[2][0] * d[0][1] - d[0][0] * d[2][1]) * tmp;
    x[2][2] = (d[0][0] * d[1][1] - d[0][1] * d[1][0]) * tmp;
    WdetJ[e] = qwt / tmp;
    for (int a = 0; a < 4; ++a)
      for (int j = 0; j < 3; ++j)
        shg[e][a][j] = shgl[a][0] * x[0][j] + shgl[a][1] * x[1][j] + shgl[a][2] * x[2][j];
  }
}

/* e3gijd, tet branch with the invariance correction (e3stab.f90:493-527) */
static void gijd_tet(const double x[3][3], double g[6]) {
  const double c1 = 1.259921049894873e+00;
  const double c2 = 6.299605249474365e-01;
  double t1, t2, t3;
  t1 = c1 * x[0][0] + c2 * (x[1][0] + x[2][0]);
  t2 = c1 * x[1][0] + c2 * (x[0][0] + x[2][0]);
  t3 = c1 * x[2][0] + c2 * (x[0][0] + x[1][0]);
  g[0] = x[0][0] * t1 + x[1][0] * t2 + x[2][0] * t3;
  t1 = c1 * x[0][1] + c2 * (x[1][1] + x[2][1]);
  t2 = c1 * x[1][1] + c2 * (x[0][1] + x[2][1]);
  t3 = c1 * x[2][1] + c2 * (x[0][1] + x[1][1]);
  g[1] = x[0][1] * t1 + x[1][1] * t2 + x[2][1] * t3;
  g[3] = x[0][0] * t1 + x[1][0] * t2 + x[2][0] * t3;
  t1 = c1 * x[0][2] + c2 * (x[1][2] + x[2][2]);
  t2 = c1 * x[1][2] + c2 * (x[0][2] + x[2][2]);
  t3 = c1 * x[2][2] + c2 * (x[0][2] + x[1][2]);
  g[2] = x[0][2] * t1 + x[1][2] * t2 + x[2][2] * t3;
  g[4] = x[0][1] * t1 + x[1][1] * t2 + x[2][1] * t3;
  g[5] = x[0][0] * t1 + x[1][0] * t2 + x[2][0] * t3;
}

void cfo_e3_block(int n, const double *yl, const double *acl,
                  const double *xl, const double *ql, double *rl,
                  double *xKebe, double *xGoC, const cfo_params *p) {
  double shp[4][4], shgl[4][3], qwt;
  tet_tables(shp, shgl, &qwt);
  double (*dxidx)[3][3] = malloc(sizeof(double[3][3]) * n);
  double (*shg)[4][3] = malloc(sizeof(double[4][3]) * n);
  double *WdetJ = malloc(sizeof(double) * n);
  const double Dtgl = 1.0 / p->Delt; /* itrPC.f90:71 */

  for (int q = 0; q < 4; ++q) { /* e3.f90:96 */
    const double *N = NULL;
    double Nq[4];
    for (int a = 0; a < 4; ++a) Nq[a] = shp[a][q]; /* getshp */
    N = Nq;
    metric(n, xl, shgl, qwt, dxidx, WdetJ, shg);
    for (int e = 0; e < n; ++e) {
      /* getdiff.f90:44-46 (Newtonian, single fluid) */
      const double rho = p->rho;
      double rmu = p->mu;
      /* e3ivar.f90:104-169: primitive variables and accelerations */
      double pres = 0.0, u1 = 0.0, u2 = 0.0, u3 = 0.0;
      for (int a = 0; a < 4; ++a) {
        pres = pres + N[a] * YL(e, a, 0);
        u1 = u1 + N[a] * YL(e, a, 1);
        u2 = u2 + N[a] * YL(e, a, 2);
        u3 = u3 + N[a] * YL(e, a, 3);
      }
      double aci[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < 4; ++a) {
        aci[0] = aci[0] + N[a] * ACL(e, a, 1);
        aci[1] = aci[1] + N[a] * ACL(e, a, 2);
        aci[2] = aci[2] + N[a] * ACL(e, a, 3);
      }
      /* mesh velocity is zero (no ALE): uMesh_i = 0 */
      const double um1 = 0.0, um2 = 0.0, um3 = 0.0;
      const double (*sg)[3] = shg[e];
      /* global gradients g1yi..g3yi (e3ivar.f90:180-198) */
      double gy[3][4];
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 4; ++k) gy[j][k] = 0.0;
      for (int a = 0; a < 4; ++a)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 4; ++k) gy[j][k] = gy[j][k] + sg[a][j] * YL(e, a, k);
      /* divergence of the reconstructed flux (e3ivar.f90:355-367) */
      double divq[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < 4; ++a) {
        divq[0] = divq[0] + sg[a][0] * QL(e, a, 0) + sg[a][1] * QL(e, a, 3) + sg[a][2] * QL(e, a, 6);
        divq[1] = divq[1] + sg[a][0] * QL(e, a, 1) + sg[a][1] * QL(e, a, 4) + sg[a][2] * QL(e, a, 7);
        divq[2] = divq[2] + sg[a][0] * QL(e, a, 2) + sg[a][1] * QL(e, a, 5) + sg[a][2] * QL(e, a, 8);
      }
      /* e3resStrongPDE_ALE (e3res.f90:775-789), src = 0 */
      const double src[3] = {0.0, 0.0, 0.0};
      double rL[3];
      rL[0] = (aci[0] + (u1 - um1) * gy[0][1] + (u2 - um2) * gy[1][1] + (u3 - um3) * gy[2][1] - src[0]) * rho + gy[0][0] - divq[0];
      rL[1] = (aci[1] + (u1 - um1) * gy[0][2] + (u2 - um2) * gy[1][2] + (u3 - um3) * gy[2][2] - src[1]) * rho + gy[1][0] - divq[1];
      rL[2] = (aci[2] + (u1 - um1) * gy[0][3] + (u2 - um2) * gy[1][3] + (u3 - um3) * gy[2][3] - src[2]) * rho + gy[2][0] - divq[2];

      /* e3stab, itau = 0 (e3stab.f90:84-133,356-374) */
      double g[6];
      gijd_tet(dxidx[e], g);
      const double fff = 36.0; /* ipord == 1 */
      const double omegasq = 0.0;
      const double rhoinv = 1.0 / rho;
      const double rnu = rmu * rhoinv;
      const double dts = Dtgl * p->dtsfct;
      const double v1 = u1 - um1, v2 = u2 - um2, v3 = u3 - um3;
      double tauM = ((2.0 * dts) * (2.0 * dts) +
                     (v1 * (g[0] * v1 + g[3] * v2 + g[5] * v3) +
                      v2 * (g[3] * v1 + g[1] * v2 + g[4] * v3) +
                      v3 * (g[5] * v1 + g[4] * v2 + g[2] * v3))) +
                    fff * (rnu * rnu) *
                        (g[0] * g[0] + g[1] * g[1] + g[2] * g[2] +
                         2. * (g[3] * g[3] + g[4] * g[4] + g[5] * g[5]) + omegasq);
      const double fact = sqrt(tauM);
      const double ff = p->taucfct / p->dtsfct;
      double tauC = rho * 0.125 * fact / (g[0] + g[1] + g[2]) * ff;
      tauM = 1.0 / fact;
      double tauBar = rL[0] * (g[0] * rL[0] + g[3] * rL[1] + g[5] * rL[2]) +
                      rL[1] * (g[3] * rL[0] + g[1] * rL[1] + g[4] * rL[2]) +
                      rL[2] * (g[5] * rL[0] + g[4] * rL[1] + g[2] * rL[2]);
      if (tauBar != 0.0) tauBar = tauM / sqrt(tauBar);
      double uBar[3];
      uBar[0] = u1 - tauM * rL[0] * rhoinv;
      uBar[1] = u2 - tauM * rL[1] * rhoinv;
      uBar[2] = u3 - tauM * rL[2] * rhoinv;

      /* e3Res, iconvflow = 2 (e3res.f90:99-286) */
      const double tmps = 1.0 - p->flmpr;
      double rNa[3], rGNa[3][3];
      rNa[0] = aci[0] * tmps - src[0];
      rNa[1] = aci[1] * tmps - src[1];
      rNa[2] = aci[2] * tmps - src[2];
      double tmp = -pres + tauC * (gy[0][1] + gy[1][2] + gy[2][3]);
      double tmp1 = rmu * (gy[1][1] + gy[0][2]);
      double tmp2 = rmu * (gy[2][2] + gy[1][3]);
      double tmp3 = rmu * (gy[0][3] + gy[2][1]);
      rNa[0] = rNa[0] + (uBar[0] - um1) * gy[0][1] + (uBar[1] - um2) * gy[1][1] + (uBar[2] - um3) * gy[2][1];
      rNa[1] = rNa[1] + (uBar[0] - um1) * gy[0][2] + (uBar[1] - um2) * gy[1][2] + (uBar[2] - um3) * gy[2][2];
      rNa[2] = rNa[2] + (uBar[0] - um1) * gy[0][3] + (uBar[1] - um2) * gy[1][3] + (uBar[2] - um3) * gy[2][3];
      rGNa[0][0] = 2.0 * rmu * gy[0][1] + tmp;
      rGNa[0][1] = tmp1;
      rGNa[0][2] = tmp3;
      rGNa[1][0] = tmp1;
      rGNa[1][1] = 2.0 * rmu * gy[1][2] + tmp;
      rGNa[1][2] = tmp2;
      rGNa[2][0] = tmp3;
      rGNa[2][1] = tmp2;
      rGNa[2][2] = 2.0 * rmu * gy[2][3] + tmp;
      tmp1 = tauM * rL[0];
      tmp2 = tauM * rL[1];
      tmp3 = tauM * rL[2];
      const double vv[3] = {u1 - um1, u2 - um2, u3 - um3};
      for (int j = 0; j < 3; ++j) {
        rGNa[0][j] = rGNa[0][j] + tmp1 * vv[j];
        rGNa[1][j] = rGNa[1][j] + tmp2 * vv[j];
        rGNa[2][j] = rGNa[2][j] + tmp3 * vv[j];
      }
      tmp1 = tauBar * (rL[0] * gy[0][1] + rL[1] * gy[1][1] + rL[2] * gy[2][1]);
      tmp2 = tauBar * (rL[0] * gy[0][2] + rL[1] * gy[1][2] + rL[2] * gy[2][2]);
      tmp3 = tauBar * (rL[0] * gy[0][3] + rL[1] * gy[1][3] + rL[2] * gy[2][3]);
      for (int j = 0; j < 3; ++j) {
        rGNa[0][j] = rGNa[0][j] + tmp1 * rL[j];
        rGNa[1][j] = rGNa[1][j] + tmp2 * rL[j];
        rGNa[2][j] = rGNa[2][j] + tmp3 * rL[j];
      }
      rNa[0] = rNa[0] * rho;
      rNa[1] = rNa[1] * rho;
      rNa[2] = rNa[2] * rho;
      const double W = WdetJ[e];
      for (int a = 0; a < 4; ++a) {
        RL(e, a, 3) = RL(e, a, 3) + W * (sg[a][0] * uBar[0] + sg[a][1] * uBar[1] + sg[a][2] * uBar[2]);
        for (int i = 0; i < 3; ++i)
          RL(e, a, i) = RL(e, a, i) - W * (N[a] * rNa[i] + sg[a][0] * rGNa[i][0] + sg[a][1] * rGNa[i][1] + sg[a][2] * rGNa[i][2]);
      }

      if (p->lhs != 1) continue;
      /* e3LHS (e3lhs.f90:85-402) */
      const double lhsFct = p->alfi * p->gami * p->Delt;
      const double lhmFct = p->almi * (1.0 - p->flmpl);
      const double tlW = lhsFct * W;
      const double t1w = tlW * rho;
      const double multfact = t1w;
      tauM = tlW * tauM;
      tauC = tlW * tauC;
      rmu = tlW * rmu;
      const double tsFct = lhmFct * W * rho;
#ifndef CFO_PLANT_LAZY_TANGENT
      /* iconvflow = 2 branch (e3lhs.f90:124-129): tauBar scaled by
         lhsFct*WdetJ and the convection taken with rho*tlW*uBar */
      tauBar = lhsFct * W * tauBar;
      double ub[3] = {t1w * uBar[0], t1w * uBar[1], t1w * uBar[2]};
#else
      /* a planted mistake, built only by the tests that show the tangent
         pins reject it: the other branch (e3lhs.f90:130-134) */
      tauBar = 0.0;
      double ub[3] = {t1w * u1, t1w * u2, t1w * u3};
#endif
      for (int b = 0; b < 4; ++b) {
        const double t1 = (ub[0] - multfact * um1) * sg[b][0] + (ub[1] - multfact * um2) * sg[b][1] + (ub[2] - multfact * um3) * sg[b][2];
        for (int a = 0; a < 4; ++a) {
          const double m1 = tsFct * N[a] * N[b];
          const double m2 = m1 + t1 * N[a];
          KE(e, 0, a, b) = KE(e, 0, a, b) + m2;
          KE(e, 4, a, b) = KE(e, 4, a, b) + m2;
          KE(e, 8, a, b) = KE(e, 8, a, b) + m2;
        }
      }
      for (int b = 0; b < 4; ++b) {
        double t1[3], t2[3], t3[3];
        for (int i = 0; i < 3; ++i) {
          t1[i] = tauC * sg[b][i];
          t2[i] = rmu * sg[b][i];
        }
        const double s1 = tauM * ((u1 - um1) * sg[b][0] + (u2 - um2) * sg[b][1] + (u3 - um3) * sg[b][2]) * rho;
        const double s2 = tauBar * (rL[0] * sg[b][0] + rL[1] * sg[b][1] + rL[2] * sg[b][2]);
        t3[0] = t2[0] + s1 * (u1 - um1) + s2 * rL[0];
        t3[1] = t2[1] + s1 * (u2 - um2) + s2 * rL[1];
        t3[2] = t2[2] + s1 * (u3 - um3) + s2 * rL[2];
        /* diagonal nodal block a == b */
        {
          const int a = b;
          const double tt = t3[0] * sg[a][0] + t3[1] * sg[a][1] + t3[2] * sg[a][2];
          KE(e, 0, a, b) = KE(e, 0, a, b) + tt + t1[0] * sg[a][0] + t2[0] * sg[a][0];
          KE(e, 4, a, b) = KE(e, 4, a, b) + tt + t1[1] * sg[a][1] + t2[1] * sg[a][1];
          KE(e, 8, a, b) = KE(e, 8, a, b) + tt + t1[2] * sg[a][2] + t2[2] * sg[a][2];
          double q1 = t1[0] * sg[a][1] + t2[1] * sg[a][0];
          KE(e, 1, a, b) = KE(e, 1, a, b) + q1;
          KE(e, 3, b, a) = KE(e, 3, b, a) + q1;
          q1 = t1[0] * sg[a][2] + t2[2] * sg[a][0];
          KE(e, 2, a, b) = KE(e, 2, a, b) + q1;
          KE(e, 6, b, a) = KE(e, 6, b, a) + q1;
          q1 = t1[1] * sg[a][2] + t2[2] * sg[a][1];
          KE(e, 5, a, b) = KE(e, 5, a, b) + q1;
          KE(e, 7, b, a) = KE(e, 7, b, a) + q1;
        }
        /* off-diagonal nodal blocks a > b */
        for (int a = b + 1; a < 4; ++a) {
          const double tt = t3[0] * sg[a][0] + t3[1] * sg[a][1] + t3[2] * sg[a][2];
          double q1 = tt + t1[0] * sg[a][0] + t2[0] * sg[a][0];
          KE(e, 0, a, b) = KE(e, 0, a, b) + q1;
          KE(e, 0, b, a) = KE(e, 0, b, a) + q1;
          q1 = tt + t1[1] * sg[a][1] + t2[1] * sg[a][1];
          KE(e, 4, a, b) = KE(e, 4, a, b) + q1;
          KE(e, 4, b, a) = KE(e, 4, b, a) + q1;
          q1 = tt + t1[2] * sg[a][2] + t2[2] * sg[a][2];
          KE(e, 8, a, b) = KE(e, 8, a, b) + q1;
          KE(e, 8, b, a) = KE(e, 8, b, a) + q1;
          q1 = t1[0] * sg[a][1] + t2[1] * sg[a][0];
          KE(e, 1, a, b) = KE(e, 1, a, b) + q1;
          KE(e, 3, b, a) = KE(e, 3, b, a) + q1;
          q1 = t1[0] * sg[a][2] + t2[2] * sg[a][0];
          KE(e, 2, a, b) = KE(e, 2, a, b) + q1;
          KE(e, 6, b, a) = KE(e, 6, b, a) + q1;
          q1 = t1[1] * sg[a][0] + t2[0] * sg[a][1];
          KE(e, 3, a, b) = KE(e, 3, a, b) + q1;
          KE(e, 1, b, a) = KE(e, 1, b, a) + q1;
          q1 = t1[1] * sg[a][2] + t2[2] * sg[a][1];
          KE(e, 5, a, b) = KE(e, 5, a, b) + q1;
          KE(e, 7, b, a) = KE(e, 7, b, a) + q1;
          q1 = t1[2] * sg[a][0] + t2[0] * sg[a][2];
          KE(e, 6, a, b) = KE(e, 6, a, b) + q1;
          KE(e, 2, b, a) = KE(e, 2, b, a) + q1;
          q1 = t1[2] * sg[a][1] + t2[1] * sg[a][2];
          KE(e, 7, a, b) = KE(e, 7, a, b) + q1;
          KE(e, 5, b, a) = KE(e, 5, b, a) + q1;
        }
      }
      /* G: xGoC(i,a,b) += tlW * N_{b,i} * N_a (e3lhs.f90:377-386) */
      for (int b = 0; b < 4; ++b) {
        const double t1[3] = {tlW * sg[b][0], tlW * sg[b][1], tlW * sg[b][2]};
        for (int a = 0; a < 4; ++a)
          for (int i = 0; i < 3; ++i) GC(e, i, a, b) = GC(e, i, a, b) + t1[i] * N[a];
      }
      /* C, lower triangle (e3lhs.f90:391-402) */
      tauM = tauM / rho;
      for (int b = 0; b < 4; ++b) {
        const double t1[3] = {tauM * sg[b][0], tauM * sg[b][1], tauM * sg[b][2]};
        for (int a = b; a < 4; ++a)
          GC(e, 3, a, b) = GC(e, 3, a, b) + t1[0] * sg[a][0] + t1[1] * sg[a][1] + t1[2] * sg[a][2];
      }
    }
  }
  /* symmetrize C (e3.f90:163-169) */
  if (p->lhs == 1)
    for (int b = 0; b < 4; ++b)
      for (int a = 0; a < b; ++a)
        for (int e = 0; e < n; ++e) GC(e, 3, a, b) = GC(e, 3, b, a);
  free(dxidx);
  free(shg);
  free(WdetJ);
}

void cfo_e3q_block(int n, const double *yl, const double *xl, double *ql,
                   double *rmassl, const cfo_params *p) {
  double shp[4][4], shgl[4][3], qwt;
  tet_tables(shp, shgl, &qwt);
  double (*dxidx)[3][3] = malloc(sizeof(double[3][3]) * n);
  double (*shg)[4][3] = malloc(sizeof(double[4][3]) * n);
  double *WdetJ = malloc(sizeof(double) * n);
  for (int q = 0; q < 4; ++q) { /* e3q.f90:66-206 */
    /* e3qvar.f90:50-140: same metric as e3metric */
    metric(n, xl, shgl, qwt, dxidx, WdetJ, shg);
    for (int e = 0; e < n; ++e) {
      const double (*sg)[3] = shg[e];
      double gy[3][4];
      for (int j = 0; j < 3; ++j)
        for (int k = 1; k < 4; ++k) gy[j][k] = 0.0;
      for (int a = 0; a < 4; ++a)
        for (int j = 0; j < 3; ++j)
          for (int k = 1; k < 4; ++k) gy[j][k] = gy[j][k] + sg[a][j] * YL(e, a, k);
      const double rmu = p->mu; /* getdiff */
      double qdi[9];
      qdi[0] = 2.0 * rmu * gy[0][1];
      qdi[3] = rmu * (gy[0][2] + gy[1][1]);
      qdi[6] = rmu * (gy[0][3] + gy[2][1]);
      qdi[1] = rmu * (gy[0][2] + gy[1][1]);
      qdi[4] = 2.0 * rmu * gy[1][2];
      qdi[7] = rmu * (gy[1][3] + gy[2][2]);
      qdi[2] = rmu * (gy[0][3] + gy[2][1]);
      qdi[5] = rmu * (gy[1][3] + gy[2][2]);
      qdi[8] = 2.0 * rmu * gy[2][3];
      const double W = WdetJ[e];
      for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 9; ++k) QL(e, a, k) = QL(e, a, k) + shp[a][q] * W * qdi[k];
      for (int a = 0; a < 4; ++a) /* row-sum lumping, idiff == 1 */
        rmassl[e + n * a] = rmassl[e + n * a] + shp[a][q] * W;
    }
  }
  free(dxidx);
  free(shg);
  free(WdetJ);
}

/* local(... 'scatter ') order: component, local node, element
 * (local.f90:68-75) */
static void scatter(int n, const int *const *ienb, int ncomp,
                    const double *loc, double *glob, int nglob) {
  for (int k = 0; k < ncomp; ++k)
    for (int a = 0; a < 4; ++a)
      for (int e = 0; e < n; ++e) {
        const int g = ienb[a][e] - 1;
        glob[g + (size_t)nglob * k] = glob[g + (size_t)nglob * k] + loc[e + n * (a + 4 * k)];
      }
}

/* sparseloc (fillsparse.f90:210-246): binary search, returns 1-based
 * position in list(1:n) */
static int sparseloc(const int *list, int n, int target) {
  int lo = 1, hi = n + 1;
  while (hi - lo > 1) {
    const int mid = (hi + lo) / 2;
    if (list[mid - 1] > target)
      hi = mid;
    else
      lo = mid;
  }
  return lo;
}

static int check_bc_supported(int nshg, const int *iBC, const double *BC, int ndofBC) {
  if (!iBC) return CFO_OK;
  for (int i = 0; i < nshg; ++i) {
    const int vb = (iBC[i] >> 3) & 7;
    /* bc3lhs projection: needs BC(:,4:6) */
    if (vb != 0 && vb != 7 && (!BC || ndofBC < 6)) return CFO_EUNSUPPORTED;
    if ((iBC[i] >> 10) & 1) return CFO_EUNSUPPORTED; /* periodic (qpbc) */
  }
  return CFO_OK;
}

/* bc3lhs (bc3lhs.f90:41-690) on one element's tangent: for each local node
 * inod with a 1- or 2-component velocity code, in local order, first the
 * column operations on the blocks (i, inod), then the row operations on the
 * blocks (inod, i), i = 1..4, then the entries of the constrained
 * components on the diagonal block (inod, inod) set to one.  Entries are
 * numbered m = 1..9 as in the reference's block-9 drawing (1 2 3 / 4 5 6 /
 * 7 8 9); each operation is entry t -= BC(in,4) * entry s1 [- BC(in,6 or 5)
 * * entry s2], exactly as the reference writes it for each code (its code-6
 * row operations included).  K(m, a, b) = xKebe[off + stride * ((m-1) + 9 *
 * (a + 4 b))]. */
typedef struct {
  signed char t, s1, s2; /* target, first source, second source (0: none), 1-based */
  signed char b1, b2;    /* BC columns (1-based) of the two sources */
} bc_op;
static void bc_apply(double *K, const bc_op *ops, int nops, const double *bcv) {
  for (int o = 0; o < nops; ++o) {
    double v = K[ops[o].t - 1] - bcv[ops[o].b1 - 1] * K[ops[o].s1 - 1];
    if (ops[o].s2) v = v - bcv[ops[o].b2 - 1] * K[ops[o].s2 - 1];
    K[ops[o].t - 1] = v;
  }
}
/* per code: column operations, zeroed entries of the columns, row
 * operations, zeroed entries of the rows, entries set to one */
typedef struct {
  bc_op col[6];
  int ncol;
  signed char colz[6];
  bc_op row[6];
  int nrow;
  signed char rowz[6];
  signed char one[2];
} bc_code;
static const bc_code BC_CODES[7] = {
    {{{0}}, 0, {0}, {{0}}, 0, {0}, {0}},
    /* 1: x1 (bc3lhs.f90:56-200) */
    {{{2, 1, 0, 4, 0}, {5, 4, 0, 4, 0}, {8, 7, 0, 4, 0}, {3, 1, 0, 5, 0}, {6, 4, 0, 5, 0}, {9, 7, 0, 5, 0}},
     6, {1, 4, 7, 0, 0, 0},
     {{4, 1, 0, 4, 0}, {5, 2, 0, 4, 0}, {6, 3, 0, 4, 0}, {7, 1, 0, 5, 0}, {8, 2, 0, 5, 0}, {9, 3, 0, 5, 0}},
     6, {1, 2, 3, 0, 0, 0}, {1, 0}},
    /* 2: x2 (:202-297) */
    {{{1, 2, 0, 4, 0}, {4, 5, 0, 4, 0}, {7, 8, 0, 4, 0}, {3, 2, 0, 5, 0}, {6, 5, 0, 5, 0}, {9, 8, 0, 5, 0}},
     6, {2, 5, 8, 0, 0, 0},
     {{1, 4, 0, 4, 0}, {2, 5, 0, 4, 0}, {3, 6, 0, 4, 0}, {7, 4, 0, 5, 0}, {8, 5, 0, 5, 0}, {9, 6, 0, 5, 0}},
     6, {4, 5, 6, 0, 0, 0}, {5, 0}},
    /* 3: x1 and x2 (:394-501) */
    {{{3, 1, 2, 4, 6}, {6, 4, 5, 4, 6}, {9, 7, 8, 4, 6}},
     3, {1, 4, 7, 2, 5, 8},
     {{7, 1, 4, 4, 6}, {8, 2, 5, 4, 6}, {9, 3, 6, 4, 6}},
     3, {1, 2, 3, 4, 5, 6}, {1, 5}},
    /* 4: x3 (:299-392) */
    {{{1, 3, 0, 4, 0}, {4, 6, 0, 4, 0}, {7, 9, 0, 4, 0}, {2, 3, 0, 5, 0}, {5, 6, 0, 5, 0}, {8, 9, 0, 5, 0}},
     6, {3, 6, 9, 0, 0, 0},
     {{1, 7, 0, 4, 0}, {2, 8, 0, 4, 0}, {3, 9, 0, 4, 0}, {4, 7, 0, 5, 0}, {5, 8, 0, 5, 0}, {6, 9, 0, 5, 0}},
     6, {7, 8, 9, 0, 0, 0}, {9, 0}},
    /* 5: x1 and x3 (:503-589) */
    {{{2, 1, 3, 4, 6}, {5, 4, 6, 4, 6}, {8, 7, 9, 4, 6}},
     3, {1, 4, 7, 3, 6, 9},
     {{4, 1, 7, 4, 6}, {5, 2, 8, 4, 6}, {6, 3, 9, 4, 6}},
     3, {1, 2, 3, 7, 8, 9}, {1, 9}},
    /* 6: x2 and x3 (:591-690; the row operations as the reference writes
     * them: row 3 from row 1, and entry 9 also from entry 6) */
    {{{1, 2, 3, 4, 6}, {4, 5, 6, 4, 6}, {7, 8, 9, 4, 6}},
     3, {2, 5, 8, 3, 6, 9},
     {{7, 1, 0, 4, 0}, {8, 2, 0, 4, 0}, {9, 3, 6, 4, 6}},
     3, {1, 2, 3, 4, 5, 6}, {5, 9}},
};
static void bc3lhs_elem(int nshg, const int *iBC, const double *BC, const int nodes[4], double *xK,
                        size_t off, size_t stride) {
  for (int inod = 0; inod < 4; ++inod) {
    const int in = nodes[inod] - 1;
    const int vb = (iBC[in] >> 3) & 7;
    if (vb == 0 || vb == 7) continue; /* bc3lhs.f90:48-49 */
    const bc_code *c = &BC_CODES[vb];
    double bcv[6];
    for (int k = 0; k < 6; ++k) bcv[k] = BC[in + (size_t)nshg * k];
    double K[9];
    for (int i = 0; i < 4; ++i) { /* columns: blocks (i, inod) */
      double *b = xK + off + stride * (size_t)(9 * (i + 4 * inod));
      for (int m = 0; m < 9; ++m) K[m] = b[stride * m];
      bc_apply(K, c->col, c->ncol, bcv);
      for (int z = 0; z < 6; ++z)
        if (c->colz[z]) K[c->colz[z] - 1] = 0.0;
      for (int m = 0; m < 9; ++m) b[stride * m] = K[m];
    }
    for (int i = 0; i < 4; ++i) { /* rows: blocks (inod, i) */
      double *b = xK + off + stride * (size_t)(9 * (inod + 4 * i));
      for (int m = 0; m < 9; ++m) K[m] = b[stride * m];
      bc_apply(K, c->row, c->nrow, bcv);
      for (int z = 0; z < 6; ++z)
        if (c->rowz[z]) K[c->rowz[z] - 1] = 0.0;
      for (int m = 0; m < 9; ++m) b[stride * m] = K[m];
    }
    double *d = xK + off + stride * (size_t)(9 * (inod + 4 * inod));
    for (int o = 0; o < 2; ++o)
      if (c->one[o]) d[stride * (c->one[o] - 1)] = 1.0;
  }
}

void cfo_bc3lhs_elem(int nshg, const int *iBC, const double *BC, const int *nodes, double *xK) {
  bc3lhs_elem(nshg, iBC, BC, nodes, xK, 0, 1);
}

typedef struct {
  int nshg, numnp, nel;
  const int *ien;
  const double *y, *ac, *x;
} mesh_view;

/* gather of one block: localy (localy.f90:54-64), localx, local */
static void gather_block(const mesh_view *m, int iel, int n, const int *const *ienb_,
                         double *yl, double *acl, double *xl, const double *q,
                         double *ql) {
  for (int a = 0; a < 4; ++a)
    for (int e = 0; e < n; ++e) {
      const int g = ienb_[a][e] - 1;
      /* yl = (p, u1, u2, u3) from y = (u1, u2, u3, p) */
      for (int j = 0; j < 3; ++j) yl[e + n * (a + 4 * (j + 1))] = m->y[g + (size_t)m->nshg * j];
      yl[e + n * (a + 4 * 0)] = m->y[g + (size_t)m->nshg * 3];
      if (acl) {
        for (int j = 0; j < 3; ++j) acl[e + n * (a + 4 * (j + 1))] = m->ac[g + (size_t)m->nshg * j];
        acl[e + n * (a + 4 * 0)] = m->ac[g + (size_t)m->nshg * 3];
      }
      for (int i = 0; i < 3; ++i) xl[e + n * (a + 4 * i)] = m->x[g + (size_t)m->numnp * i];
      if (ql)
        for (int k = 0; k < 9; ++k) ql[e + n * (a + 4 * k)] = q[g + (size_t)m->nshg * k];
    }
  (void)iel;
}

static int interior_core(int nshg, int numnp, int nel, const int *ien,
                         const double *y, const double *ac, const double *x,
                         const double *qin, int do_flux, const int *iBC,
                         const double *BC, int ndofBC,
                         const int *colm, const int *rowp, int nnz,
                         double *res, double *lhsK, double *lhsP,
                         double *qres_out, const cfo_params *p) {
  if (nshg <= 0 || nel < 0 || !ien || !y || !ac || !x || !res) return CFO_EARG;
  if (p->lhs == 1 && (!colm || !rowp || !lhsK || !lhsP)) return CFO_EARG;
  int rc = check_bc_supported(nshg, iBC, BC, ndofBC);
  if (rc) return rc;
  const int bs = p->ibksiz > 0 ? p->ibksiz : 255;
  mesh_view m = {nshg, numnp, nel, ien, y, ac, x};
  double *yl = malloc(sizeof(double) * bs * 16);
  double *acl = malloc(sizeof(double) * bs * 16);
  double *xl = malloc(sizeof(double) * bs * 12);
  double *ql = malloc(sizeof(double) * bs * 36);
  double *rl = malloc(sizeof(double) * bs * 16);
  double *rmassl = malloc(sizeof(double) * bs * 4);
  double *xKebe = malloc(sizeof(double) * bs * 144);
  double *xGoC = malloc(sizeof(double) * bs * 64);
  double *qres = qres_out ? qres_out : malloc(sizeof(double) * (size_t)nshg * 9);
  double *rmass = malloc(sizeof(double) * nshg);
  const int *ienb[4];

  if (do_flux) {
    /* diffusive flux reconstruction (elmgmr.f90:82-118) */
    memset(qres, 0, sizeof(double) * (size_t)nshg * 9);
    memset(rmass, 0, sizeof(double) * nshg);
    for (int iel = 0; iel < nel; iel += bs) {
      const int n = (nel - iel < bs) ? nel - iel : bs;
      for (int a = 0; a < 4; ++a) ienb[a] = ien + (size_t)nel * a + iel;
      /* AsIq (asiq.f90:46-75) */
      gather_block(&m, iel, n, ienb, yl, NULL, xl, NULL, NULL);
      memset(ql, 0, sizeof(double) * n * 36);
      memset(rmassl, 0, sizeof(double) * n * 4);
      cfo_e3q_block(n, yl, xl, ql, rmassl, p);
      scatter(n, ienb, 9, ql, qres, nshg);
      scatter(n, ienb, 1, rmassl, rmass, nshg);
    }
    /* qpbc (qpbc.f90:37-71), single task, no periodic nodes */
    for (int i = 0; i < nshg; ++i) rmass[i] = 1.0 / rmass[i];
    for (int k = 0; k < 9; ++k)
      for (int i = 0; i < nshg; ++i) qres[i + (size_t)nshg * k] = rmass[i] * qres[i + (size_t)nshg * k];
  } else {
    memcpy(qres, qin, sizeof(double) * (size_t)nshg * 9);
  }

  /* interior elements (elmgmr.f90:122-255) */
  memset(res, 0, sizeof(double) * (size_t)nshg * 4);
  if (p->lhs == 1) {
    memset(lhsK, 0, sizeof(double) * (size_t)nnz * 9);
    memset(lhsP, 0, sizeof(double) * (size_t)nnz * 4);
  }
  for (int iel = 0; iel < nel && rc == CFO_OK; iel += bs) {
    const int n = (nel - iel < bs) ? nel - iel : bs;
    for (int a = 0; a < 4; ++a) ienb[a] = ien + (size_t)nel * a + iel;
    /* AsIGMR (asigmr.f90:66-116) */
    gather_block(&m, iel, n, ienb, yl, acl, xl, qres, ql);
    if (p->lhs == 1) {
      memset(xKebe, 0, sizeof(double) * n * 144);
      memset(xGoC, 0, sizeof(double) * n * 64);
    }
    memset(rl, 0, sizeof(double) * n * 16);
    cfo_e3_block(n, yl, acl, xl, ql, rl, xKebe, xGoC, p);
    scatter(n, ienb, 4, rl, res, nshg);
    if (p->lhs != 1) continue;
    /* bc3lhs (elmgmr.f90:223-226), then fillsparseI (fillsparse.f90:29-63) */
    if (iBC && BC)
      for (int e = 0; e < n; ++e) {
        const int nodes[4] = {ienb[0][e], ienb[1][e], ienb[2][e], ienb[3][e]};
        bc3lhs_elem(nshg, iBC, BC, nodes, xKebe, e, n);
      }
    for (int e = 0; e < n && rc == CFO_OK; ++e)
      for (int a = 0; a < 4; ++a) {
        const int i = ienb[a][e];
        const int c = colm[i - 1];
        const int len = colm[i] - c;
        for (int b = 0; b < 4; ++b) {
          const int target = ienb[b][e];
          const int k = sparseloc(rowp + (c - 1), len, target) + c - 1;
          if (rowp[k - 1] != target) {
            rc = CFO_ESEARCH;
            break;
          }
          for (int cc = 0; cc < 9; ++cc)
            lhsK[cc + 9 * (size_t)(k - 1)] = lhsK[cc + 9 * (size_t)(k - 1)] + xKebe[e + n * (cc + 9 * (a + 4 * b))];
          for (int cc = 0; cc < 4; ++cc)
            lhsP[cc + 4 * (size_t)(k - 1)] = lhsP[cc + 4 * (size_t)(k - 1)] + xGoC[e + n * (cc + 4 * (a + 4 * b))];
        }
      }
  }
  free(yl);
  free(acl);
  free(xl);
  free(ql);
  free(rl);
  free(rmassl);
  free(xKebe);
  free(xGoC);
  if (!qres_out) free(qres);
  free(rmass);
  return rc;
}

int cfo_elmgmr_interior(int nshg, int numnp, int nel, const int *ien,
                        const double *y, const double *ac, const double *x,
                        const int *iBC, const int *colm, const int *rowp,
                        int nnz, double *res, double *lhsK, double *lhsP,
                        double *qres, const cfo_params *p) {
  return interior_core(nshg, numnp, nel, ien, y, ac, x, NULL, 1, iBC, NULL, 0, colm,
                       rowp, nnz, res, lhsK, lhsP, qres, p);
}

/* ------------------------------------------------------------------ */
/* The same interior formation in two phases, for the threaded CPU baseline:
 * (1) every element block's e3q / e3 (+ bc3lhs) into per-element storage,
 * blocks independent; (2) the scatters gathered per node or CSR row over the
 * node's element incidences.  Each sum takes its terms in the order of the
 * block loop above: local(...,'scatter ') adds by block, then local node,
 * then element (local.f90:68-75), fillsparseI by element (fillsparse.f90:
 * 29-63), so both paths give the same bits.  In the serial build CFO_PFOR is
 * empty and the phases run one after the other on one core. */
struct cfo_plan {
  int nshg, nel, bs;
  int *ptr;   /* (nshg+1): incidences of node i at ptr[i] .. ptr[i+1]-1 */
  int *inc_b; /* e << 2 | a in scatter order: block, local node, element */
  int *inc_e; /* e << 2 | a in element order */
  /* per-element storage of the last formation, kept between calls:
   * 36 + 4 + 16 + 144 + 64 doubles per element, and qres(nshg,9) */
  double *store, *qres;
};

int cfo_plan_create(int nshg, int nel, const int *ien, int ibksiz, cfo_plan **out) {
  if (nshg <= 0 || nel < 0 || !ien || !out) return CFO_EARG;
  const int bs = ibksiz > 0 ? ibksiz : 255;
  cfo_plan *pl = calloc(1, sizeof(cfo_plan));
  pl->nshg = nshg;
  pl->nel = nel;
  pl->bs = bs;
  pl->ptr = calloc((size_t)nshg + 1, sizeof(int));
  pl->inc_b = malloc(sizeof(int) * ((size_t)4 * nel + 1));
  pl->inc_e = malloc(sizeof(int) * ((size_t)4 * nel + 1));
  int *fill = malloc(sizeof(int) * ((size_t)nshg + 1));
  for (size_t k = 0; k < (size_t)4 * nel; ++k) {
    const int g = ien[k];
    if (g < 1 || g > nshg) {
      free(fill);
      cfo_plan_free(pl);
      return CFO_EARG;
    }
    pl->ptr[g] += 1;
  }
  for (int i = 0; i < nshg; ++i) pl->ptr[i + 1] += pl->ptr[i];
  memcpy(fill, pl->ptr, sizeof(int) * nshg);
  for (int e = 0; e < nel; ++e)
    for (int a = 0; a < 4; ++a) {
      const int g = ien[(size_t)nel * a + e] - 1;
      pl->inc_e[fill[g]++] = e << 2 | a;
    }
  memcpy(fill, pl->ptr, sizeof(int) * nshg);
  for (int iel = 0; iel < nel; iel += bs) {
    const int n = (nel - iel < bs) ? nel - iel : bs;
    for (int a = 0; a < 4; ++a)
      for (int e = iel; e < iel + n; ++e) {
        const int g = ien[(size_t)nel * a + e] - 1;
        pl->inc_b[fill[g]++] = e << 2 | a;
      }
  }
  free(fill);
  *out = pl;
  return CFO_OK;
}

void cfo_plan_free(cfo_plan *pl) {
  if (!pl) return;
  free(pl->ptr);
  free(pl->inc_b);
  free(pl->inc_e);
  free(pl->store);
  free(pl->qres);
  free(pl);
}

int cfo_threads(int n) {
#ifdef CFO_THREADS
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int cfo_elmgmr_interior_plan(cfo_plan *pl, int nshg, int numnp, int nel, const int *ien,
                             const double *y, const double *ac, const double *x, const int *iBC,
                             const double *BC, int ndofBC, const int *colm, const int *rowp, int nnz,
                             double *res, double *lhsK, double *lhsP, const cfo_params *p) {
  if (!pl || pl->nshg != nshg || pl->nel != nel || nel <= 0 || !ien || !y || !ac || !x || !res) return CFO_EARG;
  if (pl->bs != (p->ibksiz > 0 ? p->ibksiz : 255)) return CFO_EARG;
  if (p->lhs == 1 && (!colm || !rowp || !lhsK || !lhsP)) return CFO_EARG;
  int rc = check_bc_supported(nshg, iBC, BC, ndofBC);
  if (rc) return rc;
  const int bs = pl->bs, nblk = (nel + bs - 1) / bs;
  const int lhs = p->lhs == 1;
  mesh_view m = {nshg, numnp, nel, ien, y, ac, x};
  /* per-element storage, element-major so that a node's gather reads each
   * incidence's values contiguously: qlE[e][a][9], rmE[e][a], rlE[e][a][4],
   * KE[e][a][b][9], GE[e][a][b][4] */
  cfo_plan *pw = pl; /* the plan's storage is reused by every call */
  if (!pw->store) pw->store = malloc(sizeof(double) * (size_t)nel * (36 + 4 + 16 + 144 + 64));
  if (!pw->qres) pw->qres = malloc(sizeof(double) * (size_t)nshg * 9);
  if (!pw->store || !pw->qres) return CFO_EARG;
  double *qlE = pw->store, *rmE = qlE + (size_t)nel * 36, *rlE = rmE + (size_t)nel * 4;
  double *KE = rlE + (size_t)nel * 16, *GE = KE + (size_t)nel * 144;
  double *qres = pw->qres;
  /* flux pass, per block (AsIq, asiq.f90:46-75) */
#ifdef CFO_THREADS
#pragma omp parallel
#endif
  {
    double *yl = malloc(sizeof(double) * bs * 16), *xl = malloc(sizeof(double) * bs * 12),
           *ql = malloc(sizeof(double) * bs * 36), *rm = malloc(sizeof(double) * bs * 4);
#ifdef CFO_THREADS
#pragma omp for schedule(static)
#endif
    for (int b = 0; b < nblk; ++b) {
      const int iel = b * bs, n = (nel - iel < bs) ? nel - iel : bs;
      const int *ienb[4];
      for (int a = 0; a < 4; ++a) ienb[a] = ien + (size_t)nel * a + iel;
      gather_block(&m, iel, n, ienb, yl, NULL, xl, NULL, NULL);
      memset(ql, 0, sizeof(double) * n * 36);
      memset(rm, 0, sizeof(double) * n * 4);
      cfo_e3q_block(n, yl, xl, ql, rm, p);
      for (int e = 0; e < n; ++e)
        for (int a = 0; a < 4; ++a) {
          rmE[4 * (size_t)(iel + e) + a] = rm[e + n * a];
          for (int k = 0; k < 9; ++k) qlE[36 * (size_t)(iel + e) + 9 * a + k] = ql[e + n * (a + 4 * k)];
        }
    }
    free(yl);
    free(xl);
    free(ql);
    free(rm);
  }
  /* the flux scatter and qpbc (qpbc.f90:37-71), per node */
  CFO_PFOR
  for (int i = 0; i < nshg; ++i) {
    double rmass = 0.0, sq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = pl->ptr[i]; q < pl->ptr[i + 1]; ++q) {
      const int e = pl->inc_b[q] >> 2, a = pl->inc_b[q] & 3;
      rmass = rmass + rmE[4 * (size_t)e + a];
      const double *v = qlE + 36 * (size_t)e + 9 * a;
      for (int k = 0; k < 9; ++k) sq[k] = sq[k] + v[k];
    }
    rmass = 1.0 / rmass;
    for (int k = 0; k < 9; ++k) qres[i + (size_t)nshg * k] = rmass * sq[k];
  }
  /* interior elements, per block (AsIGMR, asigmr.f90:66-116; bc3lhs) */
#ifdef CFO_THREADS
#pragma omp parallel
#endif
  {
    double *yl = malloc(sizeof(double) * bs * 16), *acl = malloc(sizeof(double) * bs * 16),
           *xl = malloc(sizeof(double) * bs * 12), *ql = malloc(sizeof(double) * bs * 36),
           *rl = malloc(sizeof(double) * bs * 16), *xK = malloc(sizeof(double) * bs * 144),
           *xG = malloc(sizeof(double) * bs * 64);
#ifdef CFO_THREADS
#pragma omp for schedule(static)
#endif
    for (int b = 0; b < nblk; ++b) {
      const int iel = b * bs, n = (nel - iel < bs) ? nel - iel : bs;
      const int *ienb[4];
      for (int a = 0; a < 4; ++a) ienb[a] = ien + (size_t)nel * a + iel;
      gather_block(&m, iel, n, ienb, yl, acl, xl, qres, ql);
      memset(rl, 0, sizeof(double) * n * 16);
      if (lhs) {
        memset(xK, 0, sizeof(double) * n * 144);
        memset(xG, 0, sizeof(double) * n * 64);
      }
      cfo_e3_block(n, yl, acl, xl, ql, rl, lhs ? xK : NULL, lhs ? xG : NULL, p);
      if (lhs && iBC && BC)
        for (int e = 0; e < n; ++e) {
          const int nodes[4] = {ienb[0][e], ienb[1][e], ienb[2][e], ienb[3][e]};
          bc3lhs_elem(nshg, iBC, BC, nodes, xK, e, n);
        }
      for (int e = 0; e < n; ++e) {
        const size_t g = (size_t)(iel + e);
        for (int a = 0; a < 4; ++a)
          for (int k = 0; k < 4; ++k) rlE[16 * g + 4 * a + k] = rl[e + n * (a + 4 * k)];
        if (!lhs) continue;
        for (int a = 0; a < 4; ++a)
          for (int bb = 0; bb < 4; ++bb) {
            for (int cc = 0; cc < 9; ++cc) KE[144 * g + 36 * a + 9 * bb + cc] = xK[e + n * (cc + 9 * (a + 4 * bb))];
            for (int cc = 0; cc < 4; ++cc) GE[64 * g + 16 * a + 4 * bb + cc] = xG[e + n * (cc + 4 * (a + 4 * bb))];
          }
      }
    }
    free(yl);
    free(acl);
    free(xl);
    free(ql);
    free(rl);
    free(xK);
    free(xG);
  }
  /* residual rows (asigmr.f90:116) and fillsparseI rows */
  int miss = 0;
  CFO_PFOR
  for (int i = 0; i < nshg; ++i) {
    double sr[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = pl->ptr[i]; q < pl->ptr[i + 1]; ++q) {
      const int e = pl->inc_b[q] >> 2, a = pl->inc_b[q] & 3;
      const double *v = rlE + 16 * (size_t)e + 4 * a;
      for (int k = 0; k < 4; ++k) sr[k] = sr[k] + v[k];
    }
    for (int k = 0; k < 4; ++k) res[i + (size_t)nshg * k] = sr[k];
    if (!lhs) continue;
    const int c = colm[i], len = colm[i + 1] - c;
    memset(lhsK + 9 * (size_t)(c - 1), 0, sizeof(double) * 9 * len);
    memset(lhsP + 4 * (size_t)(c - 1), 0, sizeof(double) * 4 * len);
    for (int q = pl->ptr[i]; q < pl->ptr[i + 1]; ++q) {
      const int e = pl->inc_e[q] >> 2, a = pl->inc_e[q] & 3;
      const double *xKa = KE + 144 * (size_t)e + 36 * a, *xGa = GE + 64 * (size_t)e + 16 * a;
      for (int b = 0; b < 4; ++b) {
        const int target = ien[(size_t)nel * b + e];
        const int k = sparseloc(rowp + (c - 1), len, target) + c - 1;
        if (rowp[k - 1] != target) {
          miss = 1;
          continue;
        }
        for (int cc = 0; cc < 9; ++cc)
          lhsK[cc + 9 * (size_t)(k - 1)] = lhsK[cc + 9 * (size_t)(k - 1)] + xKa[9 * b + cc];
        for (int cc = 0; cc < 4; ++cc)
          lhsP[cc + 4 * (size_t)(k - 1)] = lhsP[cc + 4 * (size_t)(k - 1)] + xGa[4 * b + cc];
      }
    }
  }
  return miss ? CFO_ESEARCH : CFO_OK;
}

int cfo_elmgmr_interior_bc(int nshg, int numnp, int nel, const int *ien,
                           const double *y, const double *ac, const double *x,
                           const int *iBC, const double *BC, int ndofBC,
                           const int *colm, const int *rowp, int nnz, double *res,
                           double *lhsK, double *lhsP, const cfo_params *p) {
  return interior_core(nshg, numnp, nel, ien, y, ac, x, NULL, 1, iBC, BC, ndofBC, colm,
                       rowp, nnz, res, lhsK, lhsP, NULL, p);
}

int cfo_boundary_lhs(int nshg, int nelb, const int *ienb, const int *iBC, const double *BC,
                     int ndofBC, const int *colm, const int *rowp, double *lhsK) {
  /* the boundary loop's tangent (elmgmr.f90:364-372) for rigid walls: AsBMFG
   * leaves xKebe zero (asbmfg.f90:71), bc3lhs sets the constrained diagonal
   * entries of its partial-BC nodes to one, fillsparseI adds it */
  if (!iBC || !BC || ndofBC < 6) return CFO_EARG;
  double xK[144];
  for (int e = 0; e < nelb; ++e) {
    const int nodes[4] = {ienb[e], ienb[e + (size_t)nelb], ienb[e + 2 * (size_t)nelb], ienb[e + 3 * (size_t)nelb]};
    memset(xK, 0, sizeof(xK));
    bc3lhs_elem(nshg, iBC, BC, nodes, xK, 0, 1);
    for (int a = 0; a < 4; ++a) {
      const int i = nodes[a];
      const int c = colm[i - 1];
      const int len = colm[i] - c;
      for (int b = 0; b < 4; ++b) {
        const int k = sparseloc(rowp + (c - 1), len, nodes[b]) + c - 1;
        if (rowp[k - 1] != nodes[b]) return CFO_ESEARCH;
        for (int m = 0; m < 9; ++m) lhsK[m + 9 * (size_t)(k - 1)] += xK[m + 9 * (a + 4 * b)];
      }
    }
  }
  return CFO_OK;
}

int cfo_elmgmr_interior_q(int nshg, int numnp, int nel, const int *ien,
                          const double *y, const double *ac, const double *x,
                          const double *q, const int *iBC, const int *colm,
                          const int *rowp, int nnz, double *res, double *lhsK,
                          double *lhsP, const cfo_params *p) {
  if (!q) return CFO_EARG;
  return interior_core(nshg, numnp, nel, ien, y, ac, x, q, 0, iBC, NULL, 0, colm, rowp,
                       nnz, res, lhsK, lhsP, NULL, p);
}

int cfo_bc3res(int nshg, const int *iBC, const double *BC, int ndofBC,
               double *res) {
  /* bc3res.f90:34-80 (bc3per skipped: no periodic nodes) */
  double *r1 = res, *r2 = res + nshg, *r3 = res + 2 * (size_t)nshg, *r4 = res + 3 * (size_t)nshg;
  for (int i = 0; i < nshg; ++i) {
    if ((iBC[i] >> 10) & 1) return CFO_EUNSUPPORTED;
    const double *bc = BC ? BC : NULL;
#define B(k) (bc ? bc[i + (size_t)nshg * (k)] : 0.0)
    if ((iBC[i] >> 2) & 1) r4[i] = 0.0;
    switch ((iBC[i] >> 3) & 7) {
      case 1:
        r2[i] = r2[i] - B(3) * r1[i];
        r3[i] = r3[i] - B(4) * r1[i];
        r1[i] = 0.0;
        break;
      case 2:
        r1[i] = r1[i] - B(3) * r2[i];
        r3[i] = r3[i] - B(4) * r2[i];
        r2[i] = 0.0;
        break;
      case 3:
        r3[i] = r3[i] - B(3) * r1[i] - B(5) * r2[i];
        r1[i] = 0.0;
        r2[i] = 0.0;
        break;
      case 4:
        r1[i] = r1[i] - B(3) * r3[i];
        r2[i] = r2[i] - B(4) * r3[i];
        r3[i] = 0.0;
        break;
      case 5:
        r2[i] = r2[i] - B(3) * r1[i] - B(5) * r3[i];
        r1[i] = 0.0;
        r3[i] = 0.0;
        break;
      case 6:
        r1[i] = r1[i] - B(3) * r2[i] - B(5) * r3[i];
        r2[i] = 0.0;
        r3[i] = 0.0;
        break;
      case 7:
        r1[i] = 0.0;
        r2[i] = 0.0;
        r3[i] = 0.0;
        break;
      default:
        break;
    }
#undef B
    if ((iBC[i] >> 11) & 1) {
      r1[i] = 0.0;
      r2[i] = 0.0;
      r3[i] = 0.0;
    }
  }
  (void)ndofBC;
  return CFO_OK;
}

int cfo_genadj(int nshg, int nel, const int *ien, int *colm, int *rowp,
               int maxnnz, int *nnz_out) {
  /* Asadj (asadj.f90:27-56): per node, unique neighbour list in order of
   * first appearance over elements; genadj (genadj.f90:43-71): colm prefix
   * sum and ascending sort of each row (selection sort in the reference). */
  const int cap = 64;
  int *cnt = calloc(nshg, sizeof(int));
  int *lst = malloc(sizeof(int) * (size_t)nshg * cap);
  if (!cnt || !lst) {
    free(cnt);
    free(lst);
    return CFO_EARG;
  }
  for (int e = 0; e < nel; ++e) {
    int nd[4];
    for (int a = 0; a < 4; ++a) nd[a] = ien[e + (size_t)nel * a];
    for (int j = 0; j < 4; ++j) {
      const int jn = nd[j] - 1;
      for (int k = 0; k < 4; ++k) {
        int found = 0;
        for (int l = 0; l < cnt[jn]; ++l)
          if (lst[(size_t)jn * cap + l] == nd[k]) {
            found = 1;
            break;
          }
        if (!found) {
          if (cnt[jn] >= cap) {
            free(cnt);
            free(lst);
            return CFO_ENNZ;
          }
          lst[(size_t)jn * cap + cnt[jn]++] = nd[k];
        }
      }
    }
  }
  colm[0] = 1;
  for (int i = 0; i < nshg; ++i) colm[i + 1] = colm[i] + cnt[i];
  const int nnz = colm[nshg] - 1;
  if (nnz > maxnnz) {
    free(cnt);
    free(lst);
    return CFO_ENNZ;
  }
  int icnt = 0;
  for (int i = 0; i < nshg; ++i) {
    int *t = lst + (size_t)i * cap;
    const int ncol = cnt[i];
    for (int j = 0; j < ncol; ++j) { /* minval/minloc selection */
      int mloc = 0;
      for (int l = 1; l < ncol; ++l)
        if (t[l] < t[mloc]) mloc = l;
      rowp[icnt++] = t[mloc];
      t[mloc] = 10 * nshg + 1;
    }
  }
  *nnz_out = nnz;
  free(cnt);
  free(lst);
  return CFO_OK;
}

void cfo_pack_val4(int nshg, const int *colm, const int *rowp,
                   const double *lhsK, const double *lhsP, double *Val4) {
  /* solfar.f90:212-235 */
  const int nnz = colm[nshg] - 1;
  CFO_PFOR
  for (int i = 0; i < nnz; ++i) {
    double *v = Val4 + 16 * (size_t)i;
    const double *k = lhsK + 9 * (size_t)i, *pp = lhsP + 4 * (size_t)i;
    v[0] = k[0];
    v[1] = k[1];
    v[2] = k[2];
    v[4] = k[3];
    v[5] = k[4];
    v[6] = k[5];
    v[8] = k[6];
    v[9] = k[7];
    v[10] = k[8];
    v[12] = pp[0];
    v[13] = pp[1];
    v[14] = pp[2];
    v[15] = pp[3];
    /* entries 4,8,12 (1-based) are set by the transpose search below; the
     * reference leaves them as allocated when no transposed block exists,
     * which never happens for a symmetric graph */
    v[3] = 0.0;
    v[7] = 0.0;
    v[11] = 0.0;
  }
  CFO_PFOR
  for (int i = 1; i <= nshg; ++i)
    for (int j = colm[i - 1]; j <= colm[i] - 1; ++j) {
      const int k = rowp[j - 1];
      for (int l = colm[k - 1]; l <= colm[k] - 1; ++l)
        if (rowp[l - 1] == i) {
          double *v = Val4 + 16 * (size_t)(l - 1);
          const double *pp = lhsP + 4 * (size_t)(j - 1);
          v[3] = -pp[0];
          v[7] = -pp[1];
          v[11] = -pp[2];
          break;
        }
    }
}

/* ------------------------------------------------------------------ */
/* memLS (memLS.f90), single task: COMMUV is the identity, DOT/NORM sums are
 * the sequential loops of NORMV/NORMS/DOTV/DOTS (memLS.f90:2240-2378).     */

typedef struct {
  double nS, res;
  double *val, *valM; /* (dof, nNo) */
  int coupledFlag, incFlag;
  int nNo, dof, bGrp;
  int *glob; /* 0-based */
} face_t;

struct cfo_lhs {
  int nNo, nnz, nFaces;
  int *rowS, *rowE; /* 0-based inclusive range of each row */
  int *colPtr;      /* 0-based */
  int *diagPtr;     /* 0-based */
  face_t *face;
};

typedef struct {
  double absTol, relTol, iNorm, fNorm, dB, callD;
  int suc, mItr, sD, itr;
} subls_t;

struct cfo_ls {
  subls_t GM, CG, RI;
  int LS_type, Resm, Resc;
};

int cfo_lhs_create(int nNo, int nnz, const int *rowPtr, const int *colPtr,
                   int nFaces, cfo_lhs **out) {
  /* memLS_LHS_CREATE, nTasks == 1 branch (memLS.f90:1113-1136) */
  cfo_lhs *l = calloc(1, sizeof(cfo_lhs));
  l->nNo = nNo;
  l->nnz = nnz;
  l->nFaces = nFaces;
  l->rowS = malloc(sizeof(int) * nNo);
  l->rowE = malloc(sizeof(int) * nNo);
  l->diagPtr = malloc(sizeof(int) * nNo);
  l->colPtr = malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
  l->face = calloc(nFaces > 0 ? nFaces : 1, sizeof(face_t));
  for (int i = 0; i < nnz; ++i) l->colPtr[i] = colPtr[i] - 1;
  for (int a = 0; a < nNo; ++a) {
    const int s = rowPtr[a] - 1, e = rowPtr[a + 1] - 2;
    l->diagPtr[a] = -1;
    for (int i = s; i <= e; ++i)
      if (l->colPtr[i] == a) {
        l->diagPtr[a] = i;
        break;
      }
    l->rowS[a] = s;
    l->rowE[a] = e;
    if (l->diagPtr[a] < 0) {
      cfo_lhs_free(l);
      return CFO_ENODIAG;
    }
  }
  *out = l;
  return CFO_OK;
}

void cfo_lhs_free(cfo_lhs *l) {
  if (!l) return;
  for (int f = 0; f < l->nFaces; ++f) {
    free(l->face[f].val);
    free(l->face[f].valM);
    free(l->face[f].glob);
  }
  free(l->face);
  free(l->rowS);
  free(l->rowE);
  free(l->diagPtr);
  free(l->colPtr);
  free(l);
}

int cfo_bc_create(cfo_lhs *l, int faIn, int nNo, int dof, int bcType,
                  const int *gNodes, const double *val) {
  /* memLS_BC_CREATE (memLS.f90:272-316), single task */
  if (faIn > l->nFaces || faIn < 1) return CFO_EARG;
  face_t *f = &l->face[faIn - 1];
  free(f->val);
  free(f->valM);
  free(f->glob);
  f->nNo = nNo;
  f->dof = dof;
  f->bGrp = bcType;
  f->res = 0.0;
  f->glob = malloc(sizeof(int) * (nNo > 0 ? nNo : 1));
  f->val = calloc((size_t)dof * (nNo > 0 ? nNo : 1), sizeof(double));
  f->valM = calloc((size_t)dof * (nNo > 0 ? nNo : 1), sizeof(double));
  for (int a = 0; a < nNo; ++a) f->glob[a] = gNodes[a] - 1;
  if (val) memcpy(f->val, val, sizeof(double) * (size_t)dof * nNo);
  return CFO_OK;
}

int cfo_ls_create(int lsType, int present, double relTol, double absTol,
                  int maxItr, int dimKry, const double *relTolIn,
                  const double *absTolIn, const int *maxItrIn, cfo_ls **out) {
  /* memLS_LS_CREATE (memLS.f90:1379-1442) */
  cfo_ls *ls = calloc(1, sizeof(cfo_ls));
  ls->LS_type = lsType;
  switch (lsType) {
    case CFO_LS_NS:
      ls->RI.relTol = 4e-1;
      ls->GM.relTol = 1e-2;
      ls->CG.relTol = 1e-1;
      ls->RI.mItr = 10;
      ls->GM.mItr = 3;
      ls->CG.mItr = 500;
      ls->GM.sD = 50;
      break;
    case CFO_LS_GMRES:
      ls->RI.relTol = 1e-2;
      ls->RI.mItr = 2;
      ls->RI.sD = 150;
      break;
    case CFO_LS_CG:
      ls->RI.relTol = 1e-4;
      ls->RI.mItr = 1000;
      break;
    default:
      free(ls);
      return CFO_EARG;
  }
  ls->RI.absTol = 1e-10;
  ls->GM.absTol = 1e-10;
  ls->CG.absTol = 1e-10;
  if (present & 1) ls->RI.relTol = relTol;
  if (present & 2) ls->RI.absTol = absTol;
  if (present & 4) ls->RI.mItr = maxItr;
  if (present & 8) {
    ls->RI.sD = dimKry;
    ls->GM.sD = dimKry;
  }
  if (present & 16) {
    ls->GM.relTol = relTolIn[0];
    ls->CG.relTol = relTolIn[1];
  }
  if (present & 32) {
    ls->GM.absTol = absTolIn[0];
    ls->CG.absTol = absTolIn[1];
  }
  if (present & 64) {
    ls->GM.mItr = maxItrIn[0];
    ls->CG.mItr = maxItrIn[1];
  }
  *out = ls;
  return CFO_OK;
}

void cfo_ls_free(cfo_ls *ls) { free(ls); }

void cfo_ls_stats(const cfo_ls *ls, int sub, double *o) {
  const subls_t *s = sub == 0 ? &ls->RI : (sub == 1 ? &ls->GM : &ls->CG);
  o[0] = s->iNorm;
  o[1] = s->fNorm;
  o[2] = s->itr;
  o[3] = s->suc;
  o[4] = s->dB;
  o[5] = s->callD;
  o[6] = ls->Resm;
  o[7] = ls->Resc;
}

#ifndef CFO_THREADS
static double dotv(int dof, int nNo, const double *U, const double *V) {
  double s = 0.0;
  for (int i = 0; i < nNo; ++i) {
    double t = s;
    for (int k = 0; k < dof; ++k) t = t + U[k + (size_t)dof * i] * V[k + (size_t)dof * i];
    s = t;
  }
  return s;
}
#else
/* threaded build: each thread sums a contiguous range of rows; the partial
 * sums are added in thread order (deterministic for a fixed thread count) */
static double dotv(int dof, int nNo, const double *U, const double *V) {
  enum { kMaxThreads = 1024 };
  double part[kMaxThreads];
  int nt = 1;
#pragma omp parallel
  {
    const int T = omp_get_num_threads() < kMaxThreads ? omp_get_num_threads() : kMaxThreads;
    const int t = omp_get_thread_num();
    if (t == 0) nt = T;
    if (t < T) {
      const int i0 = (int)((long long)nNo * t / T), i1 = (int)((long long)nNo * (t + 1) / T);
      double s = 0.0;
      for (int i = i0; i < i1; ++i)
        for (int k = 0; k < dof; ++k) s = s + U[k + (size_t)dof * i] * V[k + (size_t)dof * i];
      part[t] = s;
    }
  }
  double s = 0.0;
  for (int t = 0; t < nt; ++t) s = s + part[t];
  return s;
}
#endif
static double normv(int dof, int nNo, const double *U) { return sqrt(dotv(dof, nNo, U, U)); }

/* SPARMULVV (memLS.f90:2024-2092): KU(l,i) = sum_j sum_m K(l*dof+m, j) U(m) */
static void sparmulvv(int dof, const cfo_lhs *l, const double *K,
                      const double *U, double *KU) {
  memset(KU, 0, sizeof(double) * (size_t)dof * l->nNo);
  CFO_PFOR
  for (int i = 0; i < l->nNo; ++i)
    for (int j = l->rowS[i]; j <= l->rowE[i]; ++j) {
      const int col = l->colPtr[j];
      for (int r = 0; r < dof; ++r) {
        double s = KU[r + (size_t)dof * i];
        for (int c = 0; c < dof; ++c) s = s + K[(size_t)dof * dof * j + r * dof + c] * U[c + (size_t)dof * col];
        KU[r + (size_t)dof * i] = s;
      }
    }
}
/* SPARMULVS: KU(i) = sum_j K(:,j).U(:,col) */
static void sparmulvs(int dof, const cfo_lhs *l, const double *K,
                      const double *U, double *KU) {
  memset(KU, 0, sizeof(double) * l->nNo);
  CFO_PFOR
  for (int i = 0; i < l->nNo; ++i)
    for (int j = l->rowS[i]; j <= l->rowE[i]; ++j) {
      const int col = l->colPtr[j];
      double s = KU[i];
      for (int c = 0; c < dof; ++c) s = s + K[(size_t)dof * j + c] * U[c + (size_t)dof * col];
      KU[i] = s;
    }
}
/* SPARMULSV: KU(:,i) = sum_j K(:,j) U(col) */
static void sparmulsv(int dof, const cfo_lhs *l, const double *K,
                      const double *U, double *KU) {
  memset(KU, 0, sizeof(double) * (size_t)dof * l->nNo);
  CFO_PFOR
  for (int i = 0; i < l->nNo; ++i)
    for (int j = l->rowS[i]; j <= l->rowE[i]; ++j) {
      const int col = l->colPtr[j];
      for (int c = 0; c < dof; ++c) KU[c + (size_t)dof * i] = KU[c + (size_t)dof * i] + K[(size_t)dof * j + c] * U[col];
    }
}
/* SPARMULSS */
static void sparmulss(const cfo_lhs *l, const double *K, const double *U,
                      double *KU) {
  memset(KU, 0, sizeof(double) * l->nNo);
  CFO_PFOR
  for (int i = 0; i < l->nNo; ++i)
    for (int j = l->rowS[i]; j <= l->rowE[i]; ++j) KU[i] = KU[i] + K[j] * U[l->colPtr[j]];
}

enum { BCOP_ADD = 0, BCOP_PRE = 1 };

/* ADDBCMUL (memLS.f90:206-268), unshared faces */
static void addbcmul(int op, const cfo_lhs *l, int dof, const double *X,
                     double *Y) {
  for (int f = 0; f < l->nFaces; ++f) {
    const face_t *fa = &l->face[f];
    if (!fa->coupledFlag) continue;
    const double coef = (op == BCOP_ADD) ? fa->res : -fa->res / (1.0 + fa->res * fa->nS);
    const int nsd = fa->dof < dof ? fa->dof : dof;
    double S = 0.0;
    for (int a = 0; a < fa->nNo; ++a) {
      const int Ac = fa->glob[a];
      for (int i = 0; i < nsd; ++i) S = S + fa->valM[i + (size_t)fa->dof * a] * X[i + (size_t)dof * Ac];
    }
    S = coef * S;
    for (int a = 0; a < fa->nNo; ++a) {
      const int Ac = fa->glob[a];
      for (int i = 0; i < nsd; ++i) Y[i + (size_t)dof * Ac] = Y[i + (size_t)dof * Ac] + fa->valM[i + (size_t)fa->dof * a] * S;
    }
  }
}

static int any_coupled(const cfo_lhs *l) {
  for (int f = 0; f < l->nFaces; ++f)
    if (l->face[f].coupledFlag) return 1;
  return 0;
}

/* GE (memLS.f90:607-677), A(N,N) column-major */
static int ge(int N, const double *A, double *B) {
  if (N == 2) {
    const double pivot = A[0] * A[3] - A[1] * A[2];
    const double saveEl = (B[0] * A[3] - B[1] * A[2]) / pivot;
    B[1] = (B[1] * A[0] - B[0] * A[1]) / pivot;
    B[0] = saveEl;
    return CFO_OK;
  }
  double *C = malloc(sizeof(double) * N * (N + 1));
#define CC(i, j) C[(i) + (size_t)N * (j)]
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) CC(i, j) = A[i + (size_t)N * j];
  for (int i = 0; i < N; ++i) CC(i, N) = B[i];
  for (int m = 0; m < N - 1; ++m) {
    int ipv = m;
    double pivot = fabs(CC(m, m));
    for (int i = m + 1; i < N; ++i)
      if (fabs(CC(i, m)) > pivot) {
        ipv = i;
        pivot = fabs(CC(i, m));
      }
    if (pivot < 2.0 * 2.220446049250313e-16) {
      free(C);
      return CFO_ESINGULAR;
    }
    if (ipv != m)
      for (int j = m; j <= N; ++j) {
        const double s = CC(m, j);
        CC(m, j) = CC(ipv, j);
        CC(ipv, j) = s;
      }
    for (int i = m + 1; i < N; ++i) {
      const double s = CC(i, m) / CC(m, m);
      CC(i, m) = 0.0;
      for (int j = m + 1; j <= N; ++j) CC(i, j) = CC(i, j) - s * CC(m, j);
    }
  }
  for (int j = N - 1; j >= 0; --j) {
    for (int i = j + 1; i < N; ++i) CC(j, N) = CC(j, N) - CC(j, i) * CC(i, N);
    CC(j, N) = CC(j, N) / CC(j, j);
  }
  for (int i = 0; i < N; ++i) B[i] = CC(i, N);
#undef CC
  free(C);
  return CFO_OK;
}

/* GMRES (memLS.f90:681-788); Val(dof*dof,nnz) row-major blocks */
static void gmres(const cfo_lhs *l, int dof, subls_t *ls, const double *Val,
                  const double *R, double *X) {
  const int nNo = l->nNo, sD = ls->sD;
  const size_t vn = (size_t)dof * nNo;
  double *u = malloc(sizeof(double) * vn * (sD + 1));
  double *h = calloc((size_t)(sD + 1) * sD, sizeof(double));
  double *unc = malloc(sizeof(double) * vn);
  double *y = malloc(sizeof(double) * sD), *c = malloc(sizeof(double) * sD),
         *s = malloc(sizeof(double) * sD), *err = malloc(sizeof(double) * (sD + 1));
#define H(i, j) h[(i) + (size_t)(sD + 1) * (j)]
#define U(k) (u + vn * (k))
  const int coupled = any_coupled(l);
  double eps = 0.0;
  ls->suc = 0;
  memset(X, 0, sizeof(double) * vn);
  for (int lo = 1; lo <= ls->mItr; ++lo) {
    if (lo == 1) {
      memcpy(U(0), R, sizeof(double) * vn);
    } else {
      ls->itr = ls->itr + 1;
      sparmulvv(dof, l, Val, X, U(0));
      addbcmul(BCOP_ADD, l, dof, X, U(0));
      CFO_PFOR
      for (size_t k = 0; k < vn; ++k) U(0)[k] = R[k] - U(0)[k];
    }
    if (coupled) {
      memcpy(unc, U(0), sizeof(double) * vn);
      addbcmul(BCOP_PRE, l, dof, unc, U(0));
    }
    err[0] = normv(dof, nNo, U(0));
    if (lo == 1) {
      eps = err[0];
      if (eps <= ls->absTol) {
        ls->callD = 0.0;
        ls->dB = 0.0;
        goto done;
      }
      ls->iNorm = eps;
      ls->fNorm = eps;
      eps = (ls->absTol > ls->relTol * eps) ? ls->absTol : ls->relTol * eps;
    }
    ls->dB = ls->fNorm;
    CFO_PFOR
    for (size_t k = 0; k < vn; ++k) U(0)[k] = U(0)[k] / err[0];
    int i;
    for (i = 1; i <= sD; ++i) {
      ls->itr = ls->itr + 1;
      sparmulvv(dof, l, Val, U(i - 1), U(i));
      addbcmul(BCOP_ADD, l, dof, U(i - 1), U(i));
      if (coupled) {
        memcpy(unc, U(i), sizeof(double) * vn);
        addbcmul(BCOP_PRE, l, dof, unc, U(i));
      }
      for (int j = 1; j <= i; ++j) { /* modified Gram-Schmidt */
        H(j - 1, i - 1) = dotv(dof, nNo, U(i), U(j - 1));
        CFO_PFOR
        for (size_t k = 0; k < vn; ++k) U(i)[k] = U(i)[k] - H(j - 1, i - 1) * U(j - 1)[k];
      }
      H(i, i - 1) = normv(dof, nNo, U(i));
      CFO_PFOR
      for (size_t k = 0; k < vn; ++k) U(i)[k] = U(i)[k] / H(i, i - 1);
      for (int j = 1; j <= i - 1; ++j) {
        const double tmp = c[j - 1] * H(j - 1, i - 1) + s[j - 1] * H(j, i - 1);
        H(j, i - 1) = -s[j - 1] * H(j - 1, i - 1) + c[j - 1] * H(j, i - 1);
        H(j - 1, i - 1) = tmp;
      }
      const double tmp = sqrt(H(i - 1, i - 1) * H(i - 1, i - 1) + H(i, i - 1) * H(i, i - 1));
      c[i - 1] = H(i - 1, i - 1) / tmp;
      s[i - 1] = H(i, i - 1) / tmp;
      H(i - 1, i - 1) = tmp;
      H(i, i - 1) = 0.0;
      err[i] = -s[i - 1] * err[i - 1];
      err[i - 1] = c[i - 1] * err[i - 1];
      if (fabs(err[i]) < eps) {
        ls->suc = 1;
        break;
      }
    }
    if (i > sD) i = sD;
    for (int j = 0; j < i; ++j) y[j] = err[j];
    for (int j = i; j >= 1; --j) {
      for (int k = j + 1; k <= i; ++k) y[j - 1] = y[j - 1] - H(j - 1, k - 1) * y[k - 1];
      y[j - 1] = y[j - 1] / H(j - 1, j - 1);
    }
    for (int j = 1; j <= i; ++j)
      CFO_PFOR
      for (size_t k = 0; k < vn; ++k) X[k] = X[k] + U(j - 1)[k] * y[j - 1];
    ls->fNorm = fabs(err[i]);
    if (ls->suc) break;
  }
  ls->dB = 10.0 * log(ls->fNorm / ls->dB);
done:
#undef H
#undef U
  free(u);
  free(h);
  free(unc);
  free(y);
  free(c);
  free(s);
  free(err);
}

/* GMRESV (memLS.f90:792-883) */
static void gmresv(const cfo_lhs *l, int dof, subls_t *ls, const double *Val,
                   double *R) {
  const int nNo = l->nNo, sD = ls->sD;
  const size_t vn = (size_t)dof * nNo;
  double *u = malloc(sizeof(double) * vn * (sD + 1));
  double *h = calloc((size_t)(sD + 1) * sD, sizeof(double));
  double *X = calloc(vn, sizeof(double));
  double *y = malloc(sizeof(double) * sD), *c = malloc(sizeof(double) * sD),
         *s = malloc(sizeof(double) * sD), *err = malloc(sizeof(double) * (sD + 1));
#define H(i, j) h[(i) + (size_t)(sD + 1) * (j)]
#define U(k) (u + vn * (k))
  ls->suc = 0;
  double eps = normv(dof, nNo, R);
  ls->iNorm = eps;
  ls->fNorm = eps;
  eps = (ls->absTol > ls->relTol * eps) ? ls->absTol : ls->relTol * eps;
  ls->itr = 0;
  if (ls->iNorm <= ls->absTol) {
    ls->callD = 0.0;
    ls->dB = 0.0;
    goto done;
  }
  for (int lo = 1; lo <= ls->mItr; ++lo) {
    ls->dB = ls->fNorm;
    ls->itr = ls->itr + 1;
    sparmulvv(dof, l, Val, X, U(0));
    addbcmul(BCOP_ADD, l, dof, X, U(0));
    CFO_PFOR
    for (size_t k = 0; k < vn; ++k) U(0)[k] = R[k] - U(0)[k];
    err[0] = normv(dof, nNo, U(0));
    CFO_PFOR
    for (size_t k = 0; k < vn; ++k) U(0)[k] = U(0)[k] / err[0];
    int i;
    for (i = 1; i <= sD; ++i) {
      ls->itr = ls->itr + 1;
      sparmulvv(dof, l, Val, U(i - 1), U(i));
      addbcmul(BCOP_ADD, l, dof, U(i - 1), U(i));
      for (int j = 1; j <= i; ++j) {
        H(j - 1, i - 1) = dotv(dof, nNo, U(i), U(j - 1));
        CFO_PFOR
        for (size_t k = 0; k < vn; ++k) U(i)[k] = U(i)[k] - H(j - 1, i - 1) * U(j - 1)[k];
      }
      H(i, i - 1) = normv(dof, nNo, U(i));
      CFO_PFOR
      for (size_t k = 0; k < vn; ++k) U(i)[k] = U(i)[k] / H(i, i - 1);
      for (int j = 1; j <= i - 1; ++j) {
        const double tmp = c[j - 1] * H(j - 1, i - 1) + s[j - 1] * H(j, i - 1);
        H(j, i - 1) = -s[j - 1] * H(j - 1, i - 1) + c[j - 1] * H(j, i - 1);
        H(j - 1, i - 1) = tmp;
      }
      const double tmp = sqrt(H(i - 1, i - 1) * H(i - 1, i - 1) + H(i, i - 1) * H(i, i - 1));
      c[i - 1] = H(i - 1, i - 1) / tmp;
      s[i - 1] = H(i, i - 1) / tmp;
      H(i - 1, i - 1) = tmp;
      H(i, i - 1) = 0.0;
      err[i] = -s[i - 1] * err[i - 1];
      err[i - 1] = c[i - 1] * err[i - 1];
      if (fabs(err[i]) < eps) {
        ls->suc = 1;
        break;
      }
    }
    if (i > sD) i = sD;
    for (int j = 0; j < i; ++j) y[j] = err[j];
    for (int j = i; j >= 1; --j) {
      for (int k = j + 1; k <= i; ++k) y[j - 1] = y[j - 1] - H(j - 1, k - 1) * y[k - 1];
      y[j - 1] = y[j - 1] / H(j - 1, j - 1);
    }
    for (int j = 1; j <= i; ++j)
      CFO_PFOR
      for (size_t k = 0; k < vn; ++k) X[k] = X[k] + U(j - 1)[k] * y[j - 1];
    ls->fNorm = fabs(err[i]);
    if (ls->suc) break;
  }
  memcpy(R, X, sizeof(double) * vn);
  ls->dB = 10.0 * log(ls->fNorm / ls->dB);
done:
#undef H
#undef U
  free(u);
  free(h);
  free(X);
  free(y);
  free(c);
  free(s);
  free(err);
}

/* CGRAD (memLS.f90:367-430): CG on the Schur complement L - D.G */
static void cgrad(const cfo_lhs *l, int dof, subls_t *ls, const double *D,
                  const double *G, const double *L, double *R) {
  const int nNo = l->nNo;
  double *X = calloc(nNo, sizeof(double)), *P = malloc(sizeof(double) * nNo),
         *SP = malloc(sizeof(double) * nNo), *DGP = malloc(sizeof(double) * nNo),
         *GP = malloc(sizeof(double) * (size_t)dof * nNo),
         *unc = malloc(sizeof(double) * (size_t)dof * nNo);
  const int coupled = any_coupled(l);
  ls->suc = 0;
  double errO = sqrt(dotv(1, nNo, R, R));
  ls->iNorm = errO;
  double eps = (ls->absTol > ls->relTol * errO) ? ls->absTol : ls->relTol * errO;
  eps = eps * eps;
  double err = errO;
  err = err * err;
  memcpy(P, R, sizeof(double) * nNo);
  int i;
  for (i = 1; i <= ls->mItr; ++i) {
    if (err < eps) {
      ls->suc = 1;
      break;
    }
    errO = err;
    sparmulsv(dof, l, G, P, GP);
    if (coupled) {
      memcpy(unc, GP, sizeof(double) * (size_t)dof * nNo);
      addbcmul(BCOP_PRE, l, dof, unc, GP);
    }
    sparmulvs(dof, l, D, GP, DGP);
    sparmulss(l, L, P, SP);
    CFO_PFOR
    for (int k = 0; k < nNo; ++k) SP[k] = SP[k] - DGP[k];
    const double alpha = errO / dotv(1, nNo, P, SP);
    CFO_PFOR
    for (int k = 0; k < nNo; ++k) X[k] = X[k] + alpha * P[k];
    CFO_PFOR
    for (int k = 0; k < nNo; ++k) R[k] = R[k] - alpha * SP[k];
    err = sqrt(dotv(1, nNo, R, R));
    err = err * err;
    CFO_PFOR
    for (int k = 0; k < nNo; ++k) P[k] = R[k] + err / errO * P[k];
  }
  memcpy(R, X, sizeof(double) * nNo);
  ls->fNorm = sqrt(err);
  ls->dB = 5.0 * log(err / errO);
  ls->itr = ls->itr + i - 1;
  free(X);
  free(P);
  free(SP);
  free(DGP);
  free(GP);
  free(unc);
}

/* CGRADV / CGRADS (memLS.f90:434-547) */
static void cgradv(const cfo_lhs *l, int dof, subls_t *ls, const double *K,
                   double *R) {
  const int nNo = l->nNo;
  const size_t vn = (size_t)dof * nNo;
  double *P = malloc(sizeof(double) * vn), *KP = malloc(sizeof(double) * vn), *X = calloc(vn, sizeof(double));
  ls->suc = 0;
  double err = normv(dof, nNo, R);
  ls->iNorm = err;
  double eps = (ls->absTol > ls->relTol * err) ? ls->absTol : ls->relTol * err;
  eps = eps * eps;
  err = err * err;
  memcpy(P, R, sizeof(double) * vn);
  double errO = err;
  int i;
  for (i = 1; i <= ls->mItr; ++i) {
    if (err < eps) {
      ls->suc = 1;
      break;
    }
    errO = err;
    if (dof == 1)
      sparmulss(l, K, P, KP);
    else
      sparmulvv(dof, l, K, P, KP);
    const double alpha = errO / dotv(dof, nNo, P, KP);
    CFO_PFOR
    for (size_t k = 0; k < vn; ++k) X[k] = X[k] + alpha * P[k];
    CFO_PFOR
    for (size_t k = 0; k < vn; ++k) R[k] = R[k] - alpha * KP[k];
    err = normv(dof, nNo, R);
    err = err * err;
    CFO_PFOR
    for (size_t k = 0; k < vn; ++k) P[k] = R[k] + err / errO * P[k];
  }
  memcpy(R, X, sizeof(double) * vn);
  ls->itr = i - 1;
  ls->fNorm = sqrt(err);
  ls->dB = 5.0 * log(err / errO);
  if (i > ls->mItr) ls->itr = ls->mItr;
  free(P);
  free(KP);
  free(X);
}

/* NSSOLVER (memLS.f90:1460-1759) */
static int nssolver(cfo_lhs *l, cfo_ls *ls, int dof, const double *Val,
                    double *Ri) {
  const int nNo = l->nNo, nnz = l->nnz, nsd = dof - 1;
  if (nsd != 3) return CFO_EUNSUPPORTED;
  int iB = ls->RI.mItr;
  const int nB = 2 * iB;
  const size_t mn = (size_t)nsd * nNo;
  double *Rm = malloc(sizeof(double) * mn), *Rc = malloc(sizeof(double) * nNo),
         *Rmi = malloc(sizeof(double) * mn), *Rci = malloc(sizeof(double) * nNo),
         *U = malloc(sizeof(double) * mn * iB), *P = malloc(sizeof(double) * (size_t)nNo * iB),
         *MU = malloc(sizeof(double) * mn * nB), *MP = malloc(sizeof(double) * (size_t)nNo * nB),
         *A = calloc((size_t)nB * nB, sizeof(double)), *B = calloc(nB, sizeof(double)),
         *xB = calloc(nB, sizeof(double));
  CFO_PFOR
  for (int a = 0; a < nNo; ++a) {
    for (int k = 0; k < nsd; ++k) Rmi[k + (size_t)nsd * a] = Ri[k + (size_t)dof * a];
    Rci[a] = Ri[nsd + (size_t)dof * a];
  }
  memcpy(Rm, Rmi, sizeof(double) * mn);
  memcpy(Rc, Rci, sizeof(double) * nNo);
  double t1 = normv(nsd, nNo, Rm), t2 = normv(1, nNo, Rc);
  double eps = sqrt(t1 * t1 + t2 * t2);
  ls->RI.iNorm = eps;
  ls->RI.fNorm = eps;
  ls->CG.callD = 0.0;
  ls->GM.callD = 0.0;
  ls->CG.itr = 0;
  ls->GM.itr = 0;
  ls->RI.suc = 0;
  eps = (ls->RI.absTol > ls->RI.relTol * eps) ? ls->RI.absTol : ls->RI.relTol * eps;
  /* DEPART (memLS.f90:1612-1679) */
  double *mK = malloc(sizeof(double) * 9 * (size_t)nnz), *mG = malloc(sizeof(double) * 3 * (size_t)nnz),
         *mD = malloc(sizeof(double) * 3 * (size_t)nnz), *mL = malloc(sizeof(double) * (size_t)nnz),
         *Gt = calloc(3 * (size_t)nnz, sizeof(double));
  CFO_PFOR
  for (int i = 0; i < nnz; ++i) {
    const double *t = Val + 16 * (size_t)i;
    double *k = mK + 9 * (size_t)i;
    k[0] = t[0];
    k[1] = t[1];
    k[2] = t[2];
    k[3] = t[4];
    k[4] = t[5];
    k[5] = t[6];
    k[6] = t[8];
    k[7] = t[9];
    k[8] = t[10];
    mG[3 * (size_t)i + 0] = t[3];
    mG[3 * (size_t)i + 1] = t[7];
    mG[3 * (size_t)i + 2] = t[11];
    mD[3 * (size_t)i + 0] = t[12];
    mD[3 * (size_t)i + 1] = t[13];
    mD[3 * (size_t)i + 2] = t[14];
    mL[i] = t[15];
  }
  CFO_PFOR
  for (int i = 0; i < nNo; ++i)
    for (int j = l->rowS[i]; j <= l->rowE[i]; ++j) {
      const int k = l->colPtr[j];
      for (int ll = l->rowS[k]; ll <= l->rowE[k]; ++ll)
        if (l->colPtr[ll] == i) {
          for (int c = 0; c < 3; ++c) Gt[3 * (size_t)ll + c] = -mG[3 * (size_t)j + c];
          break;
        }
    }
  /* BCPRE (memLS.f90:1683-1716) */
  for (int f = 0; f < l->nFaces; ++f) {
    face_t *fa = &l->face[f];
    if (!fa->coupledFlag) continue;
    fa->nS = 0.0;
    for (int a = 0; a < fa->nNo; ++a)
      for (int i = 0; i < nsd; ++i) fa->nS = fa->nS + fa->valM[i + (size_t)fa->dof * a] * fa->valM[i + (size_t)fa->dof * a];
  }
  int rc = CFO_OK, i, iBB = 0;
  for (i = 1; i <= ls->RI.mItr; ++i) {
    iB = 2 * i - 1;
    iBB = 2 * i;
    ls->RI.dB = ls->RI.fNorm;
    double *Ui = U + mn * (i - 1), *Pi = P + (size_t)nNo * (i - 1);
    double *MUb = MU + mn * (iB - 1), *MUbb = MU + mn * (iBB - 1);
    double *MPb = MP + (size_t)nNo * (iB - 1), *MPbb = MP + (size_t)nNo * (iBB - 1);
    gmres(l, nsd, &ls->GM, mK, Rm, Ui);
    sparmulvs(nsd, l, mD, Ui, Pi);
    CFO_PFOR
    for (int k = 0; k < nNo; ++k) Pi[k] = Rc[k] - Pi[k];
    cgrad(l, nsd, &ls->CG, Gt, mG, mL, Pi);
    sparmulsv(nsd, l, mG, Pi, MUb);
    CFO_PFOR
    for (size_t k = 0; k < mn; ++k) MUbb[k] = Rm[k] - MUb[k];
    gmres(l, nsd, &ls->GM, mK, MUbb, Ui);
    sparmulvv(nsd, l, mK, Ui, MUbb);
    addbcmul(BCOP_ADD, l, nsd, Ui, MUbb);
    sparmulss(l, mL, Pi, MPb);
    sparmulvs(nsd, l, mD, Ui, MPbb);
    for (int k = iB; k <= iBB; ++k) {
      double *MUk = MU + mn * (k - 1), *MPk = MP + (size_t)nNo * (k - 1);
      for (int j = 1; j <= k - 1; ++j) {
        const double *MUj = MU + mn * (j - 1), *MPj = MP + (size_t)nNo * (j - 1);
        A[(j - 1) + (size_t)nB * (k - 1)] = dotv(nsd, nNo, MUj, MUk) + dotv(1, nNo, MPj, MPk);
        A[(k - 1) + (size_t)nB * (j - 1)] = A[(j - 1) + (size_t)nB * (k - 1)];
      }
      const double a1 = normv(nsd, nNo, MUk), a2 = normv(1, nNo, MPk);
      A[(k - 1) + (size_t)nB * (k - 1)] = a1 * a1 + a2 * a2;
      B[k - 1] = dotv(nsd, nNo, MUk, Rmi) + dotv(1, nNo, MPk, Rci);
    }
    for (int k = 0; k < iBB; ++k) xB[k] = B[k];
    double *Asub = malloc(sizeof(double) * iBB * iBB);
    for (int c = 0; c < iBB; ++c)
      for (int r = 0; r < iBB; ++r) Asub[r + (size_t)iBB * c] = A[r + (size_t)nB * c];
    rc = ge(iBB, Asub, xB);
    free(Asub);
    if (rc) break;
    double sxb = 0.0;
    for (int k = 0; k < iBB; ++k) sxb = sxb + xB[k] * B[k];
    ls->RI.fNorm = sqrt(ls->RI.iNorm * ls->RI.iNorm - sxb);
    if (ls->RI.fNorm < eps) {
      ls->RI.suc = 1;
      break;
    }
    CFO_PFOR
    for (size_t k = 0; k < mn; ++k) Rm[k] = Rmi[k] - xB[0] * MU[k];
    CFO_PFOR
    for (int k = 0; k < nNo; ++k) Rc[k] = Rci[k] - xB[0] * MP[k];
    for (int j = 2; j <= iBB; ++j) {
      const double *MUj = MU + mn * (j - 1), *MPj = MP + (size_t)nNo * (j - 1);
      CFO_PFOR
      for (size_t k = 0; k < mn; ++k) Rm[k] = Rm[k] - xB[j - 1] * MUj[k];
      CFO_PFOR
      for (int k = 0; k < nNo; ++k) Rc[k] = Rc[k] - xB[j - 1] * MPj[k];
    }
  }
  if (rc == CFO_OK) {
    if (i > ls->RI.mItr) {
      ls->RI.itr = ls->RI.mItr;
    } else {
      ls->RI.itr = i;
      CFO_PFOR
      for (int k = 0; k < nNo; ++k) Rc[k] = Rci[k] - xB[0] * MP[k];
      for (int j = 2; j <= iBB; ++j) {
        const double *MPj = MP + (size_t)nNo * (j - 1);
        CFO_PFOR
        for (int k = 0; k < nNo; ++k) Rc[k] = Rc[k] - xB[j - 1] * MPj[k];
      }
    }
    const double rcn = normv(1, nNo, Rc) / ls->RI.fNorm;
    ls->Resc = (int)lround(100.0 * rcn * rcn);
    ls->Resm = 100 - ls->Resc;
    CFO_PFOR
    for (size_t k = 0; k < mn; ++k) Rmi[k] = xB[1] * U[k];
    CFO_PFOR
    for (int k = 0; k < nNo; ++k) Rci[k] = xB[0] * P[k];
    for (int ii = 2; ii <= ls->RI.itr; ++ii) {
      const int b1 = 2 * ii - 1, b2 = 2 * ii;
      const double *Uii = U + mn * (ii - 1), *Pii = P + (size_t)nNo * (ii - 1);
      CFO_PFOR
      for (size_t k = 0; k < mn; ++k) Rmi[k] = Rmi[k] + xB[b2 - 1] * Uii[k];
      CFO_PFOR
      for (int k = 0; k < nNo; ++k) Rci[k] = Rci[k] + xB[b1 - 1] * Pii[k];
    }
    ls->RI.dB = 10.0 * log(ls->RI.fNorm / ls->RI.dB);
    CFO_PFOR
    for (int a = 0; a < nNo; ++a) {
      for (int k = 0; k < nsd; ++k) Ri[k + (size_t)dof * a] = Rmi[k + (size_t)nsd * a];
      Ri[nsd + (size_t)dof * a] = Rci[a];
    }
  }
  free(Rm);
  free(Rc);
  free(Rmi);
  free(Rci);
  free(U);
  free(P);
  free(MU);
  free(MP);
  free(A);
  free(B);
  free(xB);
  free(mK);
  free(mG);
  free(mD);
  free(mL);
  free(Gt);
  return rc;
}

/* PRECOND (memLS.f90:1763-1943) */
static void precond(cfo_lhs *l, int dof, double *Val, double *R, double *W) {
  const int nNo = l->nNo;
  CFO_PFOR
  for (int a = 0; a < nNo; ++a) {
    const int d = l->diagPtr[a];
    for (int i = 0; i < dof; ++i) W[i + (size_t)dof * a] = Val[(size_t)dof * dof * d + i * dof + i];
  }
  CFO_PFOR
  for (int a = 0; a < nNo; ++a) {
    const int d = l->diagPtr[a];
    for (int i = 0; i < dof; ++i)
      if (W[i + (size_t)dof * a] == 0.0) {
        W[i + (size_t)dof * a] = 1.0;
        Val[(size_t)dof * dof * d + i * dof + i] = 1.0;
      }
  }
  CFO_PFOR
  for (size_t k = 0; k < (size_t)dof * nNo; ++k) W[k] = 1.0 / sqrt(fabs(W[k]));
  for (int f = 0; f < l->nFaces; ++f) {
    const face_t *fa = &l->face[f];
    if (!fa->incFlag) continue;
    const int ni = fa->dof < dof ? fa->dof : dof;
    if (fa->bGrp == CFO_BC_DIR)
      for (int a = 0; a < fa->nNo; ++a) {
        const int Ac = fa->glob[a];
        for (int i = 0; i < ni; ++i) W[i + (size_t)dof * Ac] = W[i + (size_t)dof * Ac] * fa->val[i + (size_t)fa->dof * a];
      }
  }
  /* row scaling */
  CFO_PFOR
  for (int a = 0; a < nNo; ++a)
    for (int j = l->rowS[a]; j <= l->rowE[a]; ++j)
      for (int r = 0; r < dof; ++r)
        for (int c = 0; c < dof; ++c) Val[(size_t)dof * dof * j + r * dof + c] = Val[(size_t)dof * dof * j + r * dof + c] * W[r + (size_t)dof * a];
  CFO_PFOR
  for (size_t k = 0; k < (size_t)dof * nNo; ++k) R[k] = W[k] * R[k];
  /* column scaling */
  CFO_PFOR
  for (int a = 0; a < nNo; ++a)
    for (int j = l->rowS[a]; j <= l->rowE[a]; ++j) {
      const int col = l->colPtr[j];
      for (int r = 0; r < dof; ++r)
        for (int c = 0; c < dof; ++c) Val[(size_t)dof * dof * j + r * dof + c] = Val[(size_t)dof * dof * j + r * dof + c] * W[c + (size_t)dof * col];
    }
  CFO_PFOR
  for (int a = 0; a < nNo; ++a) {
    const int d = l->diagPtr[a];
    for (int i = 0; i < dof; ++i)
      if (Val[(size_t)dof * dof * d + i * dof + i] == 0.0) Val[(size_t)dof * dof * d + i * dof + i] = 1.0;
  }
  for (int f = 0; f < l->nFaces; ++f) {
    face_t *fa = &l->face[f];
    if (!fa->coupledFlag) continue;
    const int ni = fa->dof < dof ? fa->dof : dof;
    for (int a = 0; a < fa->nNo; ++a) {
      const int Ac = fa->glob[a];
      for (int i = 0; i < ni; ++i) fa->valM[i + (size_t)fa->dof * a] = fa->val[i + (size_t)fa->dof * a] * W[i + (size_t)dof * Ac];
    }
  }
}

int cfo_ls_solve(cfo_lhs *l, cfo_ls *ls, int dof, double *Ri, double *Val,
                 const int *incL, const double *res) {
  /* memLS_SOLVE (memLS.f90:1947-2020) */
  const int nNo = l->nNo;
  for (int f = 0; f < l->nFaces; ++f) {
    face_t *fa = &l->face[f];
    fa->incFlag = 1;
    if (incL && incL[f] == 0) fa->incFlag = 0;
    fa->coupledFlag = 0;
    if (!fa->incFlag) continue;
    if (fa->bGrp == CFO_BC_NEU && res && res[f] != 0.0) {
      fa->res = res[f];
      fa->coupledFlag = 1;
    }
  }
  const size_t vn = (size_t)dof * nNo;
  double *R = malloc(sizeof(double) * vn), *W = malloc(sizeof(double) * vn);
  memcpy(R, Ri, sizeof(double) * vn);
  precond(l, dof, Val, R, W);
  int rc = CFO_OK;
  switch (ls->LS_type) {
    case CFO_LS_NS:
      rc = nssolver(l, ls, dof, Val, R);
      break;
    case CFO_LS_GMRES:
      gmresv(l, dof, &ls->RI, Val, R);
      break;
    case CFO_LS_CG:
      cgradv(l, dof, &ls->RI, Val, R);
      break;
    default:
      rc = CFO_EARG;
  }
  if (rc == CFO_OK) {
    CFO_PFOR
    for (size_t k = 0; k < vn; ++k) R[k] = R[k] * W[k];
    memcpy(Ri, R, sizeof(double) * vn);
  }
  free(R);
  free(W);
  return rc;
}

double cfo_flops_per_element(void) {
  /* Counted by the counting build (flopcount.hpp; +, -, *, /, sqrt = 1 each,
   * as written, no FMA) and pinned by the oracle tests: per element
   *   e3 (4 qp: metric, e3ivar, e3stab, e3Res, e3LHS)      7208
   *   e3q (4 qp: metric, grad u, qdi, ql/rmassl)            1524
   *   flux scatter (9 q + 1 mass) x 4 nodes                   40
   *   residual scatter 4 x 4                                  16
   *   fillsparseI 16 blocks x (9 + 4)                        208
   * = 8996; besides, 10 per node (qpbc's reciprocal and 9 products) and 25
   * per block of up to 255 elements (the shape tables of e3 and e3q). */
  return 8996.0;
}

/* ------------------------------------------------------------ boundary
 * Boundary-element residual of ElmGMR's boundary loop (elmgmr.f90:325-460):
 * AsBMFG (asbmfg.f90:19-80) gathers each boundary element (localy, localx),
 * e3b (e3b.f90:103-560) integrates over its face with e3bvar's quantities
 * (e3bvar.f90), and local(res, rl, ienb, nflow, 'scatter ') adds rl.
 * P1 tets (lcsyst 1, nshlb 3, lnode = 1,2,3: the face nodes come first in
 * ienb), rigid walls (ideformwall 0, so xKebe stays zero and the boundary
 * loop adds nothing to lhsK/lhsP), no heart model (iheart 0), no Lagrange
 * multipliers, iconvflow 2, ires 1, Newtonian rmu = mu.
 * Boundary rule: symtri (symtri.c:56-60) with 1 or 3 points, weights doubled
 * (genint.f90:86); N = (r, s, t, 1-r-s-t) at the points (shpTet), so N4 = 0
 * on the face; shglb = dN/dxi / 2 (genshpb.f90:19-20).
 */
static const double TRI_A = 0.3333333333333330000000000000000000; /* symtri.c:20 */
static const double TRI_B = 0.666666666666667;                    /* symtri.c:21 */
static const double TRI_C = 0.166666666666667;                    /* symtri.c:22 */

/* The deformable-wall inputs of the boundary loop (NULL: rigid walls). */
typedef struct {
  const double *ac, *u, *uref;
  const int *tag;
  const cfo_wall *w;
  const int *iBC;
  const double *BC;
  int ndofBC;
  const int *colm, *rowp;
  double *lhsK;
} wall_ctx;

/* e3bvar's deformable-wall section at one boundary quadrature point of a
 * vessel-wall face (e3bvar.f90:462-1017), P1 tet, lnode = 1,2,3, isotropic
 * membrane (nProps = 10, no SWB/BET files), as written: the rotation to the
 * face's flat frame, the "jacobian" of the rotated face nodes (node 4's
 * rotated coordinates left zero) and its cofactor inverse, the strain-
 * displacement matrix B(5,3 nodes,3), the membrane residual
 * rlKwall = B^T D sum_m B_m (u_m + uref_m) WdetJb h and its tangent
 * LHSwall, the wall mass vdot, the tissue support usup1 and damping velsup,
 * and at the first point the ring tethering.  Adds this point's xKebe
 * (mass, support and damping on the diagonal components, then LHSwall)
 * into xK when xK is not NULL. */
static void wall_qp(const double xl[4][3], const double bn[3], const double N[3], double WdetJb,
                    const double yl[4][4], const double acl[3][3], const double ul[3][3],
                    const double urefl[3][3], const int tagl[3], int first, const cfo_wall *w,
                    double vdot[3], double velsup[3], double usup1[3], double rlKwall[3][3], double *xK) {
  const double shglb[3][4] = {{0.5, 0.0, 0.0, -0.5}, {0.0, 0.5, 0.0, -0.5}, {0.0, 0.0, 0.5, -0.5}};
  double v1[3], v2[3], v3[3], rot[3][3];
  for (int i = 0; i < 3; ++i) {
    v3[i] = bn[i];
    v1[i] = xl[1][i] - xl[0][i];
  }
  double temp = 1.0 / sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
  for (int i = 0; i < 3; ++i) v1[i] = v1[i] * temp;
  const double c1 = v3[1] * v1[2] - v1[1] * v3[2];
  const double c2 = v1[0] * v3[2] - v3[0] * v1[2];
  const double c3 = v3[0] * v1[1] - v1[0] * v3[1];
  temp = 1.0 / sqrt(c1 * c1 + c2 * c2 + c3 * c3);
  v2[0] = c1 * temp;
  v2[1] = c2 * temp;
  v2[2] = c3 * temp;
  for (int j = 0; j < 3; ++j) {
    rot[0][j] = v1[j];
    rot[1][j] = v2[j];
    rot[2][j] = v3[j];
  }
  double xr[4][3] = {{0}};
  for (int n = 0; n < 3; ++n)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) xr[n][i] = xr[n][i] + rot[i][j] * xl[n][j];
  double J[3][3] = {{0}}, D[3][3];
  for (int n = 0; n < 4; ++n)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) J[i][j] = J[i][j] + xr[n][i] * shglb[j][n];
  D[0][0] = J[1][1] * J[2][2] - J[2][1] * J[1][2];
  D[0][1] = J[2][1] * J[0][2] - J[0][1] * J[2][2];
  D[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  temp = 1.0 / (D[0][0] * J[0][0] + D[0][1] * J[1][0] + D[0][2] * J[2][0]);
  D[0][0] = D[0][0] * temp;
  D[0][1] = D[0][1] * temp;
  D[0][2] = D[0][2] * temp;
  D[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * temp;
  D[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) * temp;
  D[1][2] = (J[1][0] * J[0][2] - J[0][0] * J[1][2]) * temp;
  D[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * temp;
  D[2][1] = (J[2][0] * J[0][1] - J[0][0] * J[2][1]) * temp;
  D[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * temp;
  /* shgb_el(i,n) = sum_j dxidxb_el(j,i) shglb(j,n) */
  double sg[3][4] = {{0}};
  for (int n = 0; n < 4; ++n)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) sg[i][n] = sg[i][n] + D[j][i] * shglb[j][n];
  /* the material matrix (e3bvar.f90:576-624, isotropic) */
  const double h = w->thicknessvw;
  const double tc = 1.0 / (1.0 - w->rnuvw * w->rnuvw);
  const double swb7 = w->evw * tc * 1.0;
  const double swb8 = w->evw * tc * w->rnuvw;
  const double swb9 = w->evw * tc * 0.5 * (1.0 - w->rnuvw);
  const double swb10 = w->evw * tc * 0.5 * (1.0 - w->rnuvw) * w->rshearconstantvw;
  double Dm[5][5] = {{0}};
  Dm[0][0] = swb7;
  Dm[1][1] = swb7;
  Dm[0][1] = swb8;
  Dm[1][0] = swb8;
  Dm[2][2] = swb9;
  Dm[3][3] = swb10;
  Dm[4][4] = swb10;
  /* B(j, n, i), the strain, D strain, D B */
  double B[5][3][3], strain[5][3] = {{0}}, Ds[5][3] = {{0}}, DB[5][3][3] = {{{0}}};
  for (int n = 0; n < 3; ++n) {
    for (int i = 0; i < 3; ++i) {
      B[0][n][i] = sg[0][n] * rot[0][i];
      B[1][n][i] = sg[1][n] * rot[1][i];
      B[2][n][i] = sg[1][n] * rot[0][i] + sg[0][n] * rot[1][i];
      B[3][n][i] = sg[0][n] * rot[2][i];
      B[4][n][i] = sg[1][n] * rot[2][i];
    }
    for (int j = 0; j < 5; ++j)
      for (int i = 0; i < 3; ++i) strain[j][n] = strain[j][n] + B[j][n][i] * (ul[n][i] + urefl[n][i]);
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) Ds[i][n] = Ds[i][n] + Dm[i][j] * strain[j][n];
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) DB[i][n][k] = DB[i][n][k] + Dm[i][j] * B[j][n][k];
  }
  double LW[9][3][3] = {{{0}}}; /* LHSwall(3(i-1)+k, n, m) */
  for (int n = 0; n < 3; ++n) {
    for (int i = 0; i < 3; ++i) rlKwall[n][i] = 0.0;
    for (int m = 0; m < 3; ++m) {
      for (int j = 0; j < 5; ++j)
        for (int i = 0; i < 3; ++i) rlKwall[n][i] = rlKwall[n][i] + B[j][n][i] * Ds[j][m];
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 5; ++j)
          for (int i = 0; i < 3; ++i)
            LW[3 * i + k][n][m] = LW[3 * i + k][n][m] + B[j][n][i] * DB[j][m][k] * WdetJb * h *
                                                            w->iwallstiffactor * w->betai * w->Delt * w->Delt *
                                                            w->alfi;
    }
    for (int i = 0; i < 3; ++i) rlKwall[n][i] = rlKwall[n][i] * WdetJb * h;
  }
  /* mass, support, damping (e3bvar.f90:783-870) */
  for (int i = 0; i < 3; ++i) vdot[i] = 0.0;
  for (int n = 0; n < 3; ++n)
    for (int i = 0; i < 3; ++i) vdot[i] = vdot[i] + N[n] * acl[n][i];
  for (int i = 0; i < 3; ++i) vdot[i] = vdot[i] * h * w->rhovw;
  double wlf = w->almi * (1.0 - w->flmpl) * w->rhovw * h;
  for (int i = 0; i < 3; ++i) usup1[i] = velsup[i] = 0.0;
  if (w->iwallsupp > 0) {
    for (int n = 0; n < 3; ++n)
      for (int i = 0; i < 3; ++i) usup1[i] = usup1[i] + N[n] * ul[n][i];
    for (int i = 0; i < 3; ++i) usup1[i] = usup1[i] * w->tissSuppStiffCoeff;
    wlf = wlf + w->Delt * w->alfi * w->betai * w->Delt * w->tissSuppStiffCoeff;
  }
  if (w->iwalldamp > 0) {
    for (int n = 0; n < 3; ++n)
      for (int i = 0; i < 3; ++i) velsup[i] = velsup[i] + N[n] * yl[n][i + 1];
    for (int i = 0; i < 3; ++i) velsup[i] = velsup[i] * w->tissSuppDampCoeff;
    wlf = wlf + w->Delt * w->alfi * w->gami * w->tissSuppDampCoeff;
  }
  /* the ring tethering, once per element (e3bvar.f90:876-975) */
  if (first && (w->iringsupp > 0 || w->iringdamp > 0) && tagl) {
    int r[3], nr = 0;
    double xm[3] = {0.0, 0.0, 0.0}, rm[3] = {0.0, 0.0, 0.0};
    for (int n = 0; n < 3; ++n) {
      if (tagl[n] & 1) {
        r[nr++] = n;
        for (int i = 0; i < 3; ++i) rm[i] = rm[i] + xl[n][i];
      }
      for (int i = 0; i < 3; ++i) xm[i] = xm[i] + xl[n][i];
    }
    for (int i = 0; i < 3; ++i) xm[i] = xm[i] / 3;
    for (int i = 0; i < 3; ++i) rm[i] = rm[i] / nr;
    if (nr > 1) {
      double d[3], t2v[3], t1v[3];
      for (int i = 0; i < 3; ++i) {
        d[i] = rm[i] - xm[i];
        t2v[i] = xl[r[1]][i] - rm[i];
      }
      double nn = sqrt(t2v[0] * t2v[0] + t2v[1] * t2v[1] + t2v[2] * t2v[2]);
      for (int i = 0; i < 3; ++i) t2v[i] = t2v[i] / nn;
      const double dp = d[0] * t2v[0] + d[1] * t2v[1] + d[2] * t2v[2];
      for (int i = 0; i < 3; ++i) d[i] = d[i] - t2v[i] * dp;
      nn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int i = 0; i < 3; ++i) d[i] = d[i] / nn;
      for (int i = 0; i < 3; ++i) t2v[i] = xl[r[1]][i] - xl[r[0]][i];
      const double len = sqrt(t2v[0] * t2v[0] + t2v[1] * t2v[1] + t2v[2] * t2v[2]) / 6.0;
      for (int i = 0; i < 3; ++i) t1v[i] = t2v[i] = 0.0;
      if (w->iringsupp > 0)
        for (int i = 0; i < 3; ++i) {
          t1v[i] = t1v[i] + w->tissSuppRingStiffCoeff * ul[r[0]][i];
          t2v[i] = t2v[i] + w->tissSuppRingStiffCoeff * ul[r[1]][i];
        }
      if (w->iringdamp > 0) /* both ends take the first ring node's velocity, as written */
        for (int i = 0; i < 3; ++i) {
          t1v[i] = t1v[i] + w->tissSuppRingDampCoeff * yl[r[0]][i + 1];
          t2v[i] = t2v[i] + w->tissSuppRingDampCoeff * yl[r[0]][i + 1];
        }
      const double p1 = d[0] * t1v[0] + d[1] * t1v[1] + d[2] * t1v[2];
      const double p2 = d[0] * t2v[0] + d[1] * t2v[1] + d[2] * t2v[2];
      for (int i = 0; i < 3; ++i) {
        t1v[i] = d[i] * p1;
        t2v[i] = d[i] * p2;
      }
      for (int i = 0; i < 3; ++i) {
        rlKwall[r[0]][i] = rlKwall[r[0]][i] + len * (2.0 * t1v[i] + t2v[i]);
        rlKwall[r[1]][i] = rlKwall[r[1]][i] + len * (t1v[i] + 2.0 * t2v[i]);
      }
      double hf = 0.0;
      if (w->iringsupp > 0) hf = hf + len * w->tissSuppRingStiffCoeff * w->betai * w->Delt * w->Delt * w->alfi;
      if (w->iringdamp > 0) hf = hf + len * w->tissSuppRingDampCoeff * w->Delt * w->alfi * w->gami;
      for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
          LW[3 * i + k][r[0]][r[0]] = LW[3 * i + k][r[0]][r[0]] + d[i] * d[k] * hf * 2.0;
          LW[3 * i + k][r[0]][r[1]] = LW[3 * i + k][r[0]][r[1]] + d[i] * d[k] * hf;
          LW[3 * i + k][r[1]][r[0]] = LW[3 * i + k][r[1]][r[0]] + d[i] * d[k] * hf;
          LW[3 * i + k][r[1]][r[1]] = LW[3 * i + k][r[1]][r[1]] + d[i] * d[k] * hf * 2.0;
        }
    }
  }
  if (!xK) return;
  wlf = w->iwallmassfactor * wlf * WdetJb;
  for (int b = 0; b < 3; ++b)
    for (int a = 0; a < 3; ++a) {
      const double t = wlf * N[a] * N[b];
      double *blk = xK + 9 * (a + 4 * b);
      blk[0] = blk[0] + t;
      blk[4] = blk[4] + t;
      blk[8] = blk[8] + t;
    }
  for (int b = 0; b < 3; ++b)
    for (int a = 0; a < 3; ++a)
      for (int m = 0; m < 9; ++m) xK[m + 9 * (a + 4 * b)] = xK[m + 9 * (a + 4 * b)] + LW[m][a][b];
}

static int boundary_core(int nshg, int numnp, int nelb, const int *ienb, const int *iBCB,
                     const double *BCB, int ndBCB, const double *y, const double *x,
                     double rho, double mu, double influx, int nqb, const int *permit,
                     int nsrf, int ibksiz, double *res, double *flx, int nflx, const wall_ctx *wc) {
  double qp[3][3], qw[3];
  if (nqb == 1) {
    qp[0][0] = qp[0][1] = qp[0][2] = TRI_A;
    qw[0] = 2.0 * 1.0;
  } else if (nqb == 3) {
    const double r[3][3] = {{TRI_B, TRI_C, TRI_C}, {TRI_C, TRI_B, TRI_C}, {TRI_C, TRI_C, TRI_B}};
    for (int q = 0; q < 3; ++q) {
      for (int j = 0; j < 3; ++j) qp[q][j] = r[q][j];
      qw[q] = 2.0 * TRI_A;
    }
  } else {
    return CFO_EARG;
  }
  if (nelb < 0 || ndBCB < 2 || ibksiz < 1) return CFO_EARG;
  for (int e = 0; e < nelb; ++e)
    if ((iBCB[e] & 4) && ndBCB < 5) return CFO_EARG;
  /* shglb for tets: d(r,s,t,1-r-s-t)/d(r,s,t), halved */
  const double shglb[3][4] = {{0.5, 0.0, 0.0, -0.5}, {0.0, 0.5, 0.0, -0.5}, {0.0, 0.0, 0.5, -0.5}};
  double *rl = (double *)malloc(sizeof(double) * (size_t)ibksiz * 16);
  if (!rl) return CFO_EARG;
  for (int e0 = 0; e0 < nelb; e0 += ibksiz) {
    const int n = (nelb - e0) < ibksiz ? (nelb - e0) : ibksiz;
    for (int i = 0; i < n * 16; ++i) rl[i] = 0.0;
    for (int el = 0; el < n; ++el) {
      const int e = e0 + el;
      int g[4];
      double yl[4][4], xl[4][3]; /* yl(a, (p,u1,u2,u3)) as localy orders it */
      for (int a = 0; a < 4; ++a) {
        g[a] = ienb[e + (size_t)nelb * a] - 1;
        yl[a][0] = y[g[a] + (size_t)nshg * 3];
        for (int i = 0; i < 3; ++i) yl[a][i + 1] = y[g[a] + (size_t)nshg * i];
        for (int i = 0; i < 3; ++i) xl[a][i] = x[g[a] + (size_t)numnp * i];
      }
      /* the wall's gathers (AsBMFG: localy of ac, localx of u) and its
       * element matrix */
      double acl[3][3] = {{0}}, ul[3][3] = {{0}}, urefl[3][3] = {{0}}, xK[144];
      int tagl[3] = {0, 0, 0};
      const int wall = wc && (iBCB[e] & 16);
      if (wc) {
        for (int a = 0; a < 3; ++a)
          for (int i = 0; i < 3; ++i) {
            acl[a][i] = wc->ac[g[a] + (size_t)nshg * i];
            ul[a][i] = wc->u[g[a] + (size_t)nshg * i];
            if (wc->uref) urefl[a][i] = wc->uref[g[a] + (size_t)nshg * i];
          }
        if (wc->tag)
          for (int a = 0; a < 3; ++a) tagl[a] = wc->tag[g[a]];
        for (int k = 0; k < 144; ++k) xK[k] = 0.0;
      }
      /* e3bvar.f90: dxdxib, its inverse by cofactors, global gradients */
      double J[3][3] = {{0}}, D[3][3], gl[3][4] = {{0}}, gy[3][4];
      for (int a = 0; a < 4; ++a)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) J[i][j] = J[i][j] + xl[a][i] * shglb[j][a];
      D[0][0] = J[1][1] * J[2][2] - J[2][1] * J[1][2];
      D[0][1] = J[2][1] * J[0][2] - J[0][1] * J[2][2];
      D[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
      const double tdet = 1.0 / (D[0][0] * J[0][0] + D[0][1] * J[1][0] + D[0][2] * J[2][0]);
      D[0][0] = D[0][0] * tdet;
      D[0][1] = D[0][1] * tdet;
      D[0][2] = D[0][2] * tdet;
      D[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * tdet;
      D[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) * tdet;
      D[1][2] = (J[1][0] * J[0][2] - J[0][0] * J[1][2]) * tdet;
      D[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * tdet;
      D[2][1] = (J[2][0] * J[0][1] - J[0][0] * J[2][1]) * tdet;
      D[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * tdet;
      for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k)
          for (int c = 0; c < 4; ++c) gl[k][c] = gl[k][c] + shglb[k][a] * yl[a][c];
      /* gy[j][c] = d(yl c)/dx_j = sum_k dxidxb(k,j) gl_k */
      for (int j = 0; j < 3; ++j)
        for (int c = 1; c < 4; ++c) gy[j][c] = D[0][j] * gl[0][c] + D[1][j] * gl[1][c] + D[2][j] * gl[2][c];
      /* outward normal and the face Jacobian (e3bvar.f90, lcsyst 1) */
      double v1[3], v2[3];
      for (int i = 0; i < 3; ++i) {
        v1[i] = xl[1][i] - xl[0][i];
        v2[i] = xl[2][i] - xl[0][i];
      }
      const double t1 = v1[1] * v2[2] - v2[1] * v1[2];
      const double t2 = v2[0] * v1[2] - v1[0] * v2[2];
      const double t3 = v1[0] * v2[1] - v2[0] * v1[1];
      const double temp = 1.0 / sqrt(t1 * t1 + t2 * t2 + t3 * t3);
      const double bn[3] = {t1 * temp, t2 * temp, t3 * temp};
      const int code = iBCB[e], srf = iBCB[e + (size_t)nelb];
      for (int q = 0; q < nqb; ++q) {
        const double N[3] = {qp[q][0], qp[q][1], qp[q][2]};
        const double WdetJb = qw[q] / (4.0 * temp);
        double pres = 0.0, u[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 3; ++a) {
          pres = pres + N[a] * yl[a][0];
          for (int i = 0; i < 3; ++i) u[i] = u[i] + N[a] * yl[a][i + 1];
        }
        double unm = bn[0] * u[0] + bn[1] * u[1] + bn[2] * u[2];
        /* gy[j][c]: g(j+1)yi(:,c+1) */
        double tau[3];
        tau[0] = bn[0] * 2.0 * mu * gy[0][1] + bn[1] * (mu * (gy[1][1] + gy[0][2])) +
                 bn[2] * (mu * (gy[2][1] + gy[0][3]));
        tau[1] = bn[0] * (mu * (gy[1][1] + gy[0][2])) + bn[1] * 2.0 * mu * gy[1][2] +
                 bn[2] * (mu * (gy[2][2] + gy[1][3]));
        tau[2] = bn[0] * (mu * (gy[2][1] + gy[0][3])) + bn[1] * (mu * (gy[2][2] + gy[1][3])) +
                 bn[2] * 2.0 * mu * gy[2][3];
        const double tn = bn[0] * tau[0] + bn[1] * tau[1] + bn[2] * tau[2];
        pres = pres - tn;
        for (int i = 0; i < 3; ++i) tau[i] = tau[i] - bn[i] * tn;
        /* surface integrals by srfID (e3b.f90:216-247): area, outflow,
         * force on the fluid */
        const int iface = srf < 0 ? -srf : srf;
        if (flx && iface != 0 && iface <= nflx) {
          double *fl = flx + 5 * (size_t)iface;
          fl[0] = fl[0] + WdetJb;
          fl[1] = fl[1] - WdetJb * unm;
          for (int i = 0; i < 3; ++i) fl[2 + i] = fl[2 + i] - (tau[i] - bn[i] * pres) * WdetJb;
        }
        /* e3b.f90: influx stabilisation, then the prescribed fluxes */
        double unm_copy = 0.0;
        if ((code & 2) && unm < 0.0) unm_copy = unm;
        if (code & 1) {
          unm = 0.0;
          for (int a = 0; a < 3; ++a) unm = unm + N[a] * BCB[e + (size_t)nelb * (a + 3 * 0)];
        }
        if (code & 2) {
          const int ok = (permit == NULL || srf < 0 || srf > nsrf) ? 1 : permit[srf];
          if (ok) {
            pres = 0.0;
            for (int a = 0; a < 3; ++a) pres = pres + N[a] * BCB[e + (size_t)nelb * (a + 3 * 1)];
          }
        }
        if (code & 4) {
          for (int i = 0; i < 3; ++i) {
            tau[i] = 0.0;
            for (int a = 0; a < 3; ++a) tau[i] = tau[i] + N[a] * BCB[e + (size_t)nelb * (a + 3 * (2 + i))];
          }
        }
        if (code & 16) {
          pres = 0.0;
          tau[0] = tau[1] = tau[2] = 0.0;
        }
        double rNa[4];
        for (int i = 0; i < 3; ++i) rNa[i] = -tau[i] + bn[i] * pres;
        rNa[3] = unm;
        double rlK[3][3];
        if (wall) { /* e3b.f90:488-511 */
          double vdot[3], velsup[3], usup1[3];
          wall_qp(xl, bn, N, WdetJb, yl, acl, ul, urefl, wc->tag ? tagl : NULL, q == 0, wc->w, vdot, velsup, usup1,
                  rlK, wc->lhsK ? xK : NULL);
          for (int i = 0; i < 3; ++i) rNa[i] = rNa[i] + vdot[i];
          if (wc->w->iwalldamp == 1)
            for (int i = 0; i < 3; ++i) rNa[i] = rNa[i] + velsup[i];
          if (wc->w->iwallsupp == 1)
            for (int i = 0; i < 3; ++i) rNa[i] = rNa[i] + usup1[i];
        }
        const double rou_copy = rho * unm_copy;
        for (int i = 0; i < 3; ++i) rNa[i] = rNa[i] - influx * rou_copy * u[i];
        for (int k = 0; k < 4; ++k) rNa[k] = rNa[k] * WdetJb;
        for (int a = 0; a < 3; ++a)
          for (int k = 0; k < 4; ++k) rl[el + n * (a + 4 * k)] = rl[el + n * (a + 4 * k)] - N[a] * rNa[k];
        if (wall) /* e3b.f90:562-575 */
          for (int a = 0; a < 3; ++a)
            for (int i = 0; i < 3; ++i) rl[el + n * (a + 4 * i)] = rl[el + n * (a + 4 * i)] - rlK[a][i];
      }
      if (wc && wc->lhsK) { /* bc3lhs and fillsparseI of this element's xKebe (elmgmr.f90:425-436) */
        const int nodes[4] = {g[0] + 1, g[1] + 1, g[2] + 1, g[3] + 1};
        bc3lhs_elem(nshg, wc->iBC, wc->BC, nodes, xK, 0, 1);
        for (int a = 0; a < 4; ++a) {
          const int c = wc->colm[g[a]];
          const int len = wc->colm[g[a] + 1] - c;
          for (int b = 0; b < 4; ++b) {
            const int k = sparseloc(wc->rowp + (c - 1), len, nodes[b]) + c - 1;
            if (wc->rowp[k - 1] != nodes[b]) {
              free(rl);
              return CFO_ESEARCH;
            }
            for (int m = 0; m < 9; ++m) wc->lhsK[m + 9 * (size_t)(k - 1)] += xK[m + 9 * (a + 4 * b)];
          }
        }
      }
    }
    /* local(res, rl, ienb, nflow, 'scatter ') (local.f90:68-75) */
    for (int k = 0; k < 4; ++k)
      for (int a = 0; a < 4; ++a)
        for (int el = 0; el < n; ++el) {
          const int gg = ienb[e0 + el + (size_t)nelb * a] - 1;
          res[gg + (size_t)nshg * k] = res[gg + (size_t)nshg * k] + rl[el + n * (a + 4 * k)];
        }
  }
  free(rl);
  return CFO_OK;
}

int cfo_boundary_res(int nshg, int numnp, int nelb, const int *ienb, const int *iBCB,
                     const double *BCB, int ndBCB, const double *y, const double *x,
                     double rho, double mu, double influx, int nqb, const int *permit,
                     int nsrf, int ibksiz, double *res) {
  return boundary_core(nshg, numnp, nelb, ienb, iBCB, BCB, ndBCB, y, x, rho, mu, influx, nqb, permit, nsrf,
                       ibksiz, res, NULL, 0, NULL);
}

int cfo_boundary_res_flx(int nshg, int numnp, int nelb, const int *ienb, const int *iBCB,
                         const double *BCB, int ndBCB, const double *y, const double *x,
                         double rho, double mu, double influx, int nqb, const int *permit,
                         int nsrf, int ibksiz, double *res, double *flxid, int nflx) {
  for (size_t k = 0; k < 5 * ((size_t)nflx + 1); ++k) flxid[k] = 0.0; /* elmgmr.f90:124 */
  return boundary_core(nshg, numnp, nelb, ienb, iBCB, BCB, ndBCB, y, x, rho, mu, influx, nqb, permit, nsrf,
                       ibksiz, res, flxid, nflx, NULL);
}

int cfo_boundary_wall(int nshg, int numnp, int nelb, const int *ienb, const int *iBCB,
                      const double *BCB, int ndBCB, const double *y, const double *ac,
                      const double *u, const double *uref, const double *x, const int *nodetag,
                      double rho, double mu, double influx, int nqb, const int *permit, int nsrf,
                      int ibksiz, const cfo_wall *w, double *res, const int *iBC, const double *BC,
                      int ndofBC, const int *colm, const int *rowp, double *lhsK) {
  if (!w || !ac || !u) return CFO_EARG;
  if (lhsK && (!iBC || !BC || ndofBC < 6 || !colm || !rowp)) return CFO_EARG;
  if ((w->iringsupp > 0 || w->iringdamp > 0) && !nodetag) return CFO_EARG;
  const wall_ctx wc = {ac, u, uref, nodetag, w, iBC, BC, ndofBC, colm, rowp, lhsK};
  return boundary_core(nshg, numnp, nelb, ienb, iBCB, BCB, ndBCB, y, x, rho, mu, influx, nqb, permit, nsrf,
                       ibksiz, res, NULL, 0, &wc);
}
