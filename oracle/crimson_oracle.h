/*
 * crimson_oracle.h -- CPU restatement of CRIMSON Flowsolver's stabilized-FEM
 * formation and memLS linear-solve path (FP64), used ONLY as test
 * infrastructure: the checker in tests/, smoke() and the CPU-baseline leg of
 * bench.py.  Nothing in the product library links or calls this code.
 *
 * Every function cites the reference file:line it restates.  Array layouts
 * follow the reference exactly: Fortran column-major, 1-based node numbers in
 * ien/colm/rowp/gNodes, node-major fields y(nshg,ndof).
 *
 * Supported flag set (rejected otherwise by the callers of this oracle):
 *   P1 tets (nshl=4), 4-point rule, itau=0, iconvflow=2, idiff=1, ipord=1,
 *   Newtonian (nnwType=0), no ALE (uMesh=0), flmpl=flmpr=0, no body force,
 *   no periodic nodes (iBC bit 10), velocity BCs with 0 or 3 components.
 *
 * Pinning.  The reference holds no kernel-level golden vectors and its
 * Fortran does not build with the compilers of this image, so each function
 * is pinned by checks the mathematics fixes (tests/test_oracle_pins.py):
 * closed forms at the zero and constant states, exact P1 mass, central
 * differences of the assembled residual against the tangent, node-permutation
 * invariance, rotation equivariance, exact flux reconstruction for linear
 * velocity, brute-force dense assembly, dense Val4 packing, and direct
 * solves for the memLS solvers.  The tauBar term tauBar (L.grad u_i) L_j of
 * the momentum flux (e3res.f90:224-245, e3stab.f90:356-367) and ubar = u -
 * tauM L / rho (e3stab.f90:372-374) are pinned on shear flows whose strong
 * residual L is known in closed form (the even and odd parts of the residual
 * in a uniform acceleration).
 * Parity unpinned: the NSSOLVER iterate sequence and iteration counts (its
 * solution is pinned); bc3lhs's code-6 row operations, restated as
 * bc3lhs.f90:648-680 writes them (no projection identity holds for them).
 */
#ifndef CRIMSON_ORACLE_H
#define CRIMSON_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double rho;     /* datmat(1,1,1) */
  double mu;      /* datmat(1,2,1) */
  double Delt;    /* Delt(itseq) */
  double alfi;    /* alpha_f */
  double almi;    /* alpha_m */
  double gami;    /* gamma */
  double dtsfct;  /* Time Constant (1) */
  double taucfct; /* C factor (1) */
  double flmpl;   /* LHS lumped mass fraction (0) */
  double flmpr;   /* RHS lumped mass fraction (0) */
  int ibksiz;     /* element block size (255) */
  int lhs;        /* 1: form the tangent */
} cfo_params;

/* error codes (0 = ok) */
enum {
  CFO_OK = 0,
  CFO_EARG = 1,        /* bad argument */
  CFO_EUNSUPPORTED = 2,/* flag/BC combination outside the supported set */
  CFO_ESEARCH = 3,     /* sparseloc target missing from the row */
  CFO_ENNZ = 4,        /* nnz capacity exceeded */
  CFO_ESINGULAR = 5,   /* GE pivot below 2 eps (memLS 'Singular matrix') */
  CFO_ENODIAG = 6      /* a CSR row without its diagonal block */
};

/* generalized-alpha constants from rho_inf, itrPC.f90:16-37 */
void cfo_time_params(double rhoinf, double *almi, double *alfi, double *gami);

/* e3 over one element block (e3.f90:1-176 with getshp, getdiff, e3ivar,
 * e3metric, e3stab/e3gijd, e3Res, e3resStrongPDE_ALE, e3LHS).
 * yl(npro,4,4) in (p,u1,u2,u3) order, acl(npro,4,4), xl(npro,4,3),
 * ql(npro,4,9).  rl(npro,4,4), xKebe(npro,9,4,4), xGoC(npro,4,4,4) are
 * ACCUMULATED into (the caller zeroes them, asigmr.f90:87-96). */
void cfo_e3_block(int npro, const double *yl, const double *acl,
                  const double *xl, const double *ql, double *rl,
                  double *xKebe, double *xGoC, const cfo_params *p);

/* e3q over one block (e3q.f90:1-211, e3qvar.f90:1-148): ql(npro,4,9) and
 * rmassl(npro,4) accumulated. */
void cfo_e3q_block(int npro, const double *yl, const double *xl, double *ql,
                   double *rmassl, const cfo_params *p);

/* Interior part of ElmGMR (elmgmr.f90:82-255): AsIq block loop, qpbc,
 * zeroing, AsIGMR/e3 block loop with residual scatter, bc3lhs (no-op for the
 * supported BC codes) and fillsparseI.  ien(nel,4) 1-based; y,ac (nshg,4);
 * x (numnp,3); colm (nshg+1), rowp (nnz) 1-based.  Outputs res (nshg,4),
 * lhsK (9,nnz), lhsP (4,nnz) (lhs==1 only), qres (nshg,9) optional. */
int cfo_elmgmr_interior(int nshg, int numnp, int nel, const int *ien,
                        const double *y, const double *ac, const double *x,
                        const int *iBC, const int *colm, const int *rowp,
                        int nnz, double *res, double *lhsK, double *lhsP,
                        double *qres, const cfo_params *p);

/* Same with the diffusive flux q supplied (skips AsIq/qpbc); used to
 * differentiate the element residual with q frozen. */
int cfo_elmgmr_interior_q(int nshg, int numnp, int nel, const int *ien,
                          const double *y, const double *ac, const double *x,
                          const double *q, const int *iBC, const int *colm,
                          const int *rowp, int nnz, double *res, double *lhsK,
                          double *lhsP, const cfo_params *p);

/* cfo_boundary_res plus the boundary loop's surface integrals by srfID
 * (e3b.f90:216-247): flxid(5, 0..nflx) zeroed, then for each boundary
 * element with |srfID| in 1..nflx, at each quadrature point: area, the
 * outflow -u.n, and the force on the fluid -(tau - n p) (tau and p after
 * e3bvar's normal-traction split). */
int cfo_boundary_res_flx(int nshg, int numnp, int nelb, const int *ienb, const int *iBCB,
                         const double *BCB, int ndBCB, const double *y, const double *x,
                         double rho, double mu, double influx, int nqb, const int *permit,
                         int nsrf, int ibksiz, double *res, double *flxid, int nflx);

/* The same with BC(nshg,ndofBC) (ndofBC >= 6): 1- and 2-component velocity
 * codes are projected out of each element's tangent by bc3lhs before the
 * assembly (elmgmr.f90:223-226, bc3lhs.f90:41-690). */
int cfo_elmgmr_interior_bc(int nshg, int numnp, int nel, const int *ien,
                           const double *y, const double *ac, const double *x,
                           const int *iBC, const double *BC, int ndofBC,
                           const int *colm, const int *rowp, int nnz, double *res,
                           double *lhsK, double *lhsP, const cfo_params *p);

/* bc3lhs on one element's xKebe(9,4,4) (xK[(m-1) + 9 (a + 4 b)]); nodes
 * 1-based. */
void cfo_bc3lhs_elem(int nshg, const int *iBC, const double *BC, const int *nodes, double *xK);

/* The boundary loop's tangent for rigid walls (elmgmr.f90:364-372): a zero
 * xKebe per boundary element through bc3lhs and fillsparseI, added to
 * lhsK(9,nnz).  ienb(nelb,4) 1-based. */
int cfo_boundary_lhs(int nshg, int nelb, const int *ienb, const int *iBC, const double *BC,
                     int ndofBC, const int *colm, const int *rowp, double *lhsK);

/* Deformable-wall parameters (solver.inp "Deformable Wall" block,
 * input_fform.cxx:1237-1262; isotropic membrane, no SWB/BET files). */
typedef struct {
  double rhovw;            /* Density of Vessel Wall */
  double thicknessvw;      /* Thickness of Vessel Wall = SWB(:,1) */
  double evw;              /* Young Mod of Vessel Wall */
  double rnuvw;            /* Poisson Ratio of Vessel Wall */
  double rshearconstantvw; /* Shear Constant of Vessel Wall */
  double tissSuppStiffCoeff, tissSuppDampCoeff;         /* iwallsupp, iwalldamp */
  double tissSuppRingStiffCoeff, tissSuppRingDampCoeff; /* iringsupp, iringdamp */
  int iwallmassfactor, iwallstiffactor;                 /* Wall Mass/Stiffness Matrix for LHS */
  int iwallsupp, iwalldamp, iringsupp, iringdamp;
  double Delt, alfi, almi, gami, betai, flmpl;          /* time integration (betai = 1) */
} cfo_wall;

/* The boundary loop with deformable walls (ideformwall = 1): cfo_boundary_res
 * plus, on the vessel-wall faces (iBCB bit 4), e3bvar's membrane, mass,
 * tissue-support, damping and ring-tethering terms (e3bvar.f90:462-1017)
 * and e3b's use of them (e3b.f90:374-395,488-575).  ac(nshg,4) as y;
 * u(nshg,3) the displacement (uAlpha), uref(nshg,3) the reference
 * displacement (NULL: zero), nodetag(numnp) the node tags (bit 0: ring;
 * needed for the ring terms).  When lhsK is not NULL, each boundary
 * element's xKebe (the wall tangent, zero elsewhere) goes through bc3lhs
 * (iBC, BC(nshg,ndofBC)) and fillsparseI into lhsK(9,nnz)
 * (elmgmr.f90:425-436). */
int cfo_boundary_wall(int nshg, int numnp, int nelb, const int *ienb, const int *iBCB,
                      const double *BCB, int ndBCB, const double *y, const double *ac,
                      const double *u, const double *uref, const double *x, const int *nodetag,
                      double rho, double mu, double influx, int nqb, const int *permit, int nsrf,
                      int ibksiz, const cfo_wall *w, double *res, const int *iBC, const double *BC,
                      int ndofBC, const int *colm, const int *rowp, double *lhsK);

/* bc3Res without periodicity (bc3res.f90:1-80). BC(nshg,ndofBC). */
int cfo_bc3res(int nshg, const int *iBC, const double *BC, int ndofBC,
               double *res);

/* genadj + Asadj (genadj.f90:1-84, asadj.f90:1-60): colm(nshg+1), rowp. */
int cfo_genadj(int nshg, int nel, const int *ien, int *colm, int *rowp,
               int maxnnz, int *nnz_out);

/* Val4/Res4 packing of SolFlow (solfar.f90:206-235). */
void cfo_pack_val4(int nshg, const int *colm, const int *rowp,
                   const double *lhsK, const double *lhsP, double *Val4);

/* ---------------- memLS (memLS.f90) ---------------- */
enum { CFO_LS_CG = 1, CFO_LS_GMRES = 2, CFO_LS_NS = 3 };
enum { CFO_BC_DIR = 0, CFO_BC_NEU = 1 };

typedef struct cfo_lhs cfo_lhs;
typedef struct cfo_ls cfo_ls;

/* memLS_LHS_CREATE, single task (memLS.f90:1080-1136). rowPtr(nNo+1) and
 * colPtr(nnz) 1-based, as passed from itrdrv.f90:934 (rowPtr=colm,
 * colPtr=rowp). */
int cfo_lhs_create(int nNo, int nnz, const int *rowPtr, const int *colPtr,
                   int nFaces, cfo_lhs **out);
void cfo_lhs_free(cfo_lhs *lhs);
/* memLS_BC_CREATE (memLS.f90:272-342); val may be NULL (=> zeros). */
int cfo_bc_create(cfo_lhs *lhs, int faIn, int nNo, int dof, int bcType,
                  const int *gNodes, const double *val);
/* memLS_LS_CREATE (memLS.f90:1379-1442).  'present' bit mask:
 * 1 relTol, 2 absTol, 4 maxItr, 8 dimKry, 16 relTolIn, 32 absTolIn,
 * 64 maxItrIn. */
int cfo_ls_create(int lsType, int present, double relTol, double absTol,
                  int maxItr, int dimKry, const double *relTolIn,
                  const double *absTolIn, const int *maxItrIn, cfo_ls **out);
void cfo_ls_free(cfo_ls *ls);
/* memLS_SOLVE (memLS.f90:1947-2020). Ri(dof,nNo) in/out, Val(dof*dof,nnz)
 * in/out (rescaled in place as the reference does).  incL/res may be NULL. */
int cfo_ls_solve(cfo_lhs *lhs, cfo_ls *ls, int dof, double *Ri, double *Val,
                 const int *incL, const double *res);
/* stats: for sub = 0 (RI), 1 (GM), 2 (CG): out[0]=iNorm out[1]=fNorm
 * out[2]=itr out[3]=suc out[4]=dB out[5]=callD; plus Resm/Resc in out[6..7] */
void cfo_ls_stats(const cfo_ls *ls, int sub, double *out);

/* Boundary-element residual of ElmGMR's boundary loop (AsBMFG -> e3b ->
 * e3bvar, elmgmr.f90:325-460): res(nshg,4) += the boundary elements' rl.
 * ienb(nelb,4) 1-based, face nodes first; iBCB(nelb,2) (code bits, srfID);
 * BCB(nelb,3,ndBCB) (mass flux, pressure, traction x3); y(nshg,4) =
 * (u1,u2,u3,p); nqb = 1 or 3 boundary points; permit(0..nsrf) flow-permitted
 * flags by srfID (NULL: all); influx = the Influx Coefficient (0.5).
 * P1 tets, rigid walls, no heart model: the boundary tangent is zero. */
int cfo_boundary_res(int nshg, int numnp, int nelb, const int *ienb, const int *iBCB,
                     const double *BCB, int ndBCB, const double *y, const double *x,
                     double rho, double mu, double influx, int nqb, const int *permit,
                     int nsrf, int ibksiz, double *res);

/* Flop count of one element formation in the reference formulation
 * (flux pass + e3 with lhs=1 + scatter adds), counted stage by stage from
 * the restatement above (+,-,*,/,sqrt each 1). */
double cfo_flops_per_element(void);

/* The interior formation of cfo_elmgmr_interior_bc (without qres output)
 * in two phases over a plan of the mesh's node incidences: element blocks
 * first (e3q, then e3 + bc3lhs, into per-element storage), the scatters
 * then gathered per node and CSR row in the block loop's order, so the
 * result is bit-identical.  Built with -fopenmp -DCFO_THREADS
 * (liboracle_mt.so) both phases run on cfo_threads() threads: the all-core
 * CPU baseline.  ibksiz of the plan must equal p->ibksiz.  The plan keeps the
 * per-element storage between calls (one call at a time per plan). */
typedef struct cfo_plan cfo_plan;
int cfo_plan_create(int nshg, int nel, const int *ien, int ibksiz, cfo_plan **out);
void cfo_plan_free(cfo_plan *pl);
int cfo_elmgmr_interior_plan(cfo_plan *pl, int nshg, int numnp, int nel, const int *ien,
                             const double *y, const double *ac, const double *x, const int *iBC,
                             const double *BC, int ndofBC, const int *colm, const int *rowp, int nnz,
                             double *res, double *lhsK, double *lhsP, const cfo_params *p);
/* Threads of the threaded build (n > 0 sets them); 1 in the serial build. */
int cfo_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
