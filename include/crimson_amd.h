/*
 * crimson_amd.h -- C ABI of the MI355X-native formation and memLS solve path.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or
 * one of the CRIMSON_E* codes; crimson_last_error() returns a text for the
 * last failure on the calling thread.  Host-pointer entry points copy their
 * inputs to the GPU and their outputs back, and are synchronous at return.
 * *_dev entry points take device pointers and return once their work is
 * complete on the handle's stream.
 *
 * Index base: like the reference, ien/colm/rowp/gNodes are 1-based.
 * Layouts: Fortran column-major, y(nshg,4) columns (u1,u2,u3,p);
 * lhsK(9,nnz) and lhsP(4,nnz) as fillsparseI writes them; Val(dof*dof,nnz)
 * row-major blocks as memLS reads them; Ri(dof,nNo).
 */
#ifndef CRIMSON_AMD_H
#define CRIMSON_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CRIMSON_OK = 0,
  CRIMSON_EARG = 1,         /* invalid argument or shape */
  CRIMSON_EUNSUPPORTED = 2, /* flag or BC combination outside the supported set */
  CRIMSON_EGRAPH = 3,       /* element pair missing from the CSR graph, or graph not symmetric */
  CRIMSON_ENODIAG = 4,      /* a CSR row without its diagonal block */
  CRIMSON_EHIP = 5,         /* HIP runtime error (no device, launch failure) */
  CRIMSON_ENOMEM = 6,       /* device allocation failed */
  CRIMSON_ESINGULAR = 7,    /* memLS GE pivot below 2*eps ('Singular matrix') */
  CRIMSON_ESTATE = 8        /* handle used out of order */
};

/* Copies the text of the calling thread's last error into buf (len bytes). */
int crimson_last_error(char *buf, int len);
/* Number of HIP devices visible (0 without a GPU); never fails. */
int crimson_device_count(void);

/* ------------------------------------------------------------------ */
/* Element formation: interior part of ElmGMR.                         */

typedef struct {
  double rho;     /* datmat(1,1,1) */
  double mu;      /* datmat(1,2,1), Newtonian */
  double Delt;    /* time step */
  double alfi;    /* alpha_f */
  double almi;    /* alpha_m */
  double gami;    /* gamma */
  double dtsfct;  /* Time Constant for tau (1) */
  double taucfct; /* C factor for tauC (1) */
  double flmpl;   /* LHS lumped-mass fraction: only 0 supported */
  double flmpr;   /* RHS lumped-mass fraction: only 0 supported */
  int itau;       /* only 0 */
  int iconvflow;  /* only 2 (advective form) */
  int idiff;      /* only 1 (global flux reconstruction, row-sum lumping) */
  int ipord;      /* only 1 (P1 tets, 4-point rule) */
  int precision;  /* 64: FP64 formation; 32: FP64 residual with the
                     tangent formed and assembled in FP32 with FP32
                     accumulation (metric and tau in FP64 on elements whose
                     aspect ratio exceeds aspect_fp64) */
  double aspect_fp64; /* aspect-ratio threshold of the FP64 tau island:
                         longest edge over shortest altitude, 1 for the
                         regular tet; used when precision == 32 (>= 1) */
  int ideformwall;    /* "Deformable Wall": 0 rigid walls (vessel-wall faces,
                         iBCB bit 4, carry no traction and no membrane
                         terms); 1 the vessel-wall faces carry e3bvar's
                         membrane, mass, support and damping terms and their
                         tangent (crimson_form_set_wall, required before the
                         first formation) */
} crimson_form_params;

/* Deformable-wall values (solver.inp's wall block, input_fform.cxx:
 * 1237-1262): an isotropic membrane of uniform thickness ("Number of Wall
 * Properties per Node" 10, no SWB, TWB or BET files).  The time integration
 * (Delt, alfi, almi, gami) is the formation's, betai = 1 (itrPC.f90:31-35). */
typedef struct {
  double rhovw;            /* Density of Vessel Wall */
  double thicknessvw;      /* Thickness of Vessel Wall (SWB(:,1)) */
  double evw;              /* Young Mod of Vessel Wall */
  double rnuvw;            /* Poisson Ratio of Vessel Wall */
  double rshearconstantvw; /* Shear Constant of Vessel Wall */
  double tissSuppStiffCoeff;     /* Stiffness Coefficient for Tissue Support */
  double tissSuppDampCoeff;      /* Damping Coefficient for Tissue Support */
  double tissSuppRingStiffCoeff; /* Axial Tethering Stiffness Coefficient */
  double tissSuppRingDampCoeff;  /* Axial Tethering Damping Coefficient */
  int iwallmassfactor;     /* Wall Mass Matrix for LHS (1) */
  int iwallstiffactor;     /* Wall Stiffness Matrix for LHS (1) */
  int iwallsupp;           /* Wall External Support Term */
  int iwalldamp;           /* Wall Damping Term */
  int iringsupp;           /* Axial Tethering Stiffness Term */
  int iringdamp;           /* Axial Tethering Damping Term */
} crimson_wall_params;

typedef struct crimson_form crimson_form;

/* Replaces the setup ElmGMR relies on: the element blocks mien(iblk)%p
 * concatenated (elmgmr.f90:134-148), the CSR graph from genadj
 * (genadj.f90:1-84) and the BC codes bc3lhs reads (bc3lhs.f90:44-49).
 * ien(nel,4) 1-based column-major; colm(nshg+1)/rowp(nnz) 1-based;
 * iBC(nshg) may be NULL.  Builds the element->nonzero maps once. */
int crimson_form_setup(int nshg, int numnp, int nel, const int *ien,
                       const int *colm, const int *rowp, int nnz,
                       const int *iBC, const crimson_form_params *prm,
                       crimson_form **out);

/* ElmGMR interior (elmgmr.f90:82-255): AsIq + qpbc, AsIGMR/e3 with the
 * residual scatter, bc3lhs, fillsparseI.  Host arrays.  res(nshg,4) is
 * overwritten; lhsK(9,nnz) and lhsP(4,nnz) are overwritten when lhs == 1
 * (the reference zeroes them first, elmgmr.f90:127-130). qres(nshg,9) is an
 * optional output of the reconstructed diffusive flux (may be NULL). */
int crimson_form_interior(crimson_form *f, const double *y, const double *ac,
                          const double *x, double *res, double *lhsK,
                          double *lhsP, double *qres, int lhs);

/* Same on device pointers; inputs and outputs are FP64 device arrays in the
 * layouts above whatever the handle's formation precision. */
int crimson_form_interior_dev(crimson_form *f, const double *d_y,
                              const double *d_ac, const double *d_x,
                              double *d_res, double *d_lhsK, double *d_lhsP,
                              double *d_qres, int lhs);

/* Device-resident fused form: formation writes the memLS matrix of `lhs`
 * directly (the Val4 packing of solfar.f90:212-235, with the transposed
 * -G blocks taken from the same element contributions instead of the
 * transpose search) and the residual as d_Res4 (4,nshg), the transposed
 * copy SolFlow passes to memLS_SOLVE (solfar.f90:206-210). */
typedef struct crimson_lhs crimson_lhs;
typedef struct crimson_ls crimson_ls;
int crimson_form_interior_to_lhs(crimson_form *f, const double *d_y,
                                 const double *d_ac, const double *d_x,
                                 double *d_Res4, struct crimson_lhs *lhs);

/* bc3Res (bc3res.f90:34-120, no periodic nodes) on a device residual:
 * res(nshg,4) when res4_layout is 0, Res4(4,nshg) when it is 1.  A no-op
 * for a handle set up without iBC; 1- and 2-component velocity codes read
 * BC (crimson_form_set_bc, else CRIMSON_ESTATE). */
int crimson_form_bc3res_dev(crimson_form *f, double *d_res, int res4_layout);

/* BC(nshg, ndofBC) of the handle's iBC (column-major, the reference's
 * BC array; ndofBC >= 6 when iBC has 1- or 2-component velocity codes,
 * whose projections read BC(:,4:6)).  With it every formation with the
 * tangent applies bc3lhs (bc3lhs.f90:41-690, called by ElmGMR for the
 * interior elements at elmgmr.f90:223-226 and for the boundary elements at
 * :364-372) to the matrix, and crimson_form_bc3res_dev the partial codes'
 * residual projection.  Formations with the tangent of a handle whose iBC
 * has such codes fail with CRIMSON_ESTATE until BC is set.  ndofBC = 0
 * removes it.  (The setup rejects a code-6 node that shares an element with
 * another such node: the reference's code-6 row operations then depend on
 * the element's local node order.) */
int crimson_form_set_bc(crimson_form *f, const double *BC, int ndofBC);

/* Wall time (ms) of the last formation's kernels, measured with HIP events
 * on the handle's stream: out[0] flux pass, out[1] element kernel,
 * out[2] assembly of the interior rows, out[3] total (with the boundary
 * elements, crimson_form_set_boundary). */
int crimson_form_timing(const crimson_form *f, double *out4);

/* Boundary elements: replaces the boundary-element loop of ElmGMR
 * (elmgmr.f90:325-460: AsBMFG asbmfg.f90:1-80 -> e3b e3b.f90:1-560 -> e3bvar
 * e3bvar.f90).  After this call every formation of the handle adds the
 * boundary elements' residual rows to res / Res4 after the interior rows, as
 * the reference's boundary loop does; with rigid walls the matrix is
 * unchanged (the boundary xKebe is zero).  ienb(nelb,4) 1-based with the three face
 * nodes first (the mienb blocks concatenated), iBCB(nelb,2) (code bits: 1
 * mass flux, 2 pressure, 4 traction, 16 vessel wall; srfID), BCB(nelb,3,
 * ndBCB) (mass flux, pressure, traction x3; ndBCB >= 5 when bit 4 is used),
 * nqb = 1 or 3 boundary points ("Quadrature Rule on Boundary" 1 or 2), influx
 * = "Influx Coefficient", permit(0..nsrf) the flow-permitted flag of each
 * srfID (callCPPDiscoverWhetherFlowPermittedAcrossSurface, e3b.f90; NULL: all
 * permitted).  nelb = 0 removes the boundary (and the wall).  Deformable
 * walls come with crimson_form_set_wall; the heart model and Lagrange
 * multipliers are not formed.  With ensemble instances each boundary element
 * takes the viscosity of the instance that holds its nodes.  The VJP
 * (crimson_form_vjp_dev) of a handle with boundary elements includes their
 * rows (rigid walls only). */
int crimson_form_set_boundary(crimson_form *f, int nelb, const int *ienb,
                              const int *iBCB, const double *BCB, int ndBCB,
                              int nqb, double influx, const int *permit,
                              int nsrf);

/* Deformable walls (ideformwall = 1; vlmwStTri vlmwsttri.f90:60-210 and
 * gendat.f90:172-225 at set-up): after crimson_form_set_boundary, the faces
 * with iBCB bit 4 become the vessel wall and every formation adds, on them,
 * e3bvar's wall terms (e3bvar.f90:462-1017 through e3b.f90:374-575): the
 * membrane B^T D B (u + uref) in each face's flat frame, the wall mass
 * rhovw h ac, the tissue support k u and damping c v, the ring tethering
 * along the end rings, and their tangent into the velocity blocks of the
 * matrix (before bc3lhs, as AsBMFG's xKebe).  nodetag(numnp): the node tags
 * (bit 0 "ring"; NULL when neither ring term is on); uref(nshg,3)
 * column-major: the reference displacement ("displacement_ref", NULL:
 * zero).  The wall nodes (mWNodes) are the faces' nodes.  A face with three
 * ring nodes is refused (CRIMSON_EUNSUPPORTED: e3bvar's tethering direction
 * is then 0/0); a wall face whose plane passes through the origin gives
 * non-finite rows, as e3bvar's flat-frame inverse does.  Calling
 * crimson_form_set_boundary again removes the wall.  The VJP refuses a
 * handle with a wall; the captured timestep (crimson_step_graph_create)
 * carries the wall displacement given by crimson_form_set_wall_state_dev. */
int crimson_form_set_wall(crimson_form *f, const crimson_wall_params *w,
                          const int *nodetag, const double *uref);
/* The displacement u(nshg,3) (column-major, device) at which the following
 * formations evaluate the wall terms: the reference passes uAlpha to ElmGMR
 * (solfar.f90:148).  The pointer is kept; it must stay valid while used. */
int crimson_form_set_displacement_dev(crimson_form *f, const double *d_u);
/* The displacement at t_n and t_{n+1} (device u(nshg,3) arrays) that a
 * timestep captured by crimson_step_graph_create updates with the wall's
 * predictor, alpha level and corrector, copying u into uold at its end (as
 * it copies y and ac); with crimson_form_set_displacement_dev's uAlpha.
 * Pointers kept; no counterpart in the reference (its arrays are global). */
int crimson_form_set_wall_state_dev(crimson_form *f, double *d_uold, double *d_u);
/* The wall nodes' displacement in the generalized-alpha step, on device
 * arrays u(nshg,3), y(nshg,4), ac(nshg,4), Res4(4,nshg):
 *   predict  u = uold + Delt yold + (gami - 2 betai)/(2 gami) Delt^2 acold
 *            (itrPredict, itrPC.f90:127-132)
 *   correct  u += Delt^2 betai solinc (itrCorrect, itrPC.f90:183-186)
 *   alpha    uAlpha = uold + alfi (u - uold) (itrYAlpha, itrPC.f90:333-336)
 * Other nodes are left alone. */
int crimson_wall_predict_dev(crimson_form *f, const double *d_uold,
                             const double *d_yold, const double *d_acold,
                             double *d_u);
int crimson_wall_correct_dev(crimson_form *f, double *d_u, const double *d_Res4);
int crimson_wall_alpha_dev(crimson_form *f, const double *d_uold,
                           const double *d_u, double *d_uAlpha);

/* Autotuner (no counterpart in the reference): times the formation into
 * the memLS handle `lhs` (crimson_form_interior_to_lhs, reps calls after one
 * warm-up per candidate, assembly phase from HIP events) for each launch
 * configuration and keeps the fastest for the handle's later calls:
 * candidate c = 0..11: row schedule c % 3 (0 the caller's numbering, 1
 * Morton order of the nodes, 2 runs of 64 consecutive rows in Morton order
 * of their centroids), 8-incidence assembly chunks if (c / 3) % 2 (else
 * 16; FP64 handles only), workgroups dealt to the XCDs in turns of 64 if
 * c / 6 (else one contiguous range per XCD).  mask selects the candidates
 * tried (bit c = candidate c; 0 = all valid ones).  *choice receives the
 * winner, ms_out (12 doubles, may be NULL) the mean assembly milliseconds
 * of each candidate tried.  No configuration changes a
 * result; the matrix and res hold the last candidate's formation. */
int crimson_form_autotune(crimson_form *f, const double *d_y, const double *d_ac,
                          const double *d_x, double *d_res, crimson_lhs *lhs,
                          int reps, int mask, int *choice, double *ms_out);

/* The boundary loop's surface integrals of the last formation (flxID of
 * e3b.f90:216-247, zeroed by ElmGMR at elmgmr.f90:124): flxid(5, 0..nflx)
 * column-major; column s sums, over the boundary elements whose srfID is
 * +-s, at each quadrature point: the area, the outflow -u.n and the force
 * on the fluid -(tau - n p) (tau, p after e3bvar's normal-traction split).
 * Column 0 and surfaces without elements are zero. */
int crimson_form_boundary_flux(crimson_form *f, int nflx, double *flxid);

/* crimson_form_boundary_flux without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate (code near HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-surface sums are staged in a fixed device array of the library
 * (5 doubles per |srfID| below CRIMSON_FLUX_NOALLOC_MAX_SURFACES, 160 KiB,
 * held by each device the library runs on until the process ends) instead
 * of a buffer of the handle.  Boundary elements with |srfID| >=
 * CRIMSON_FLUX_NOALLOC_MAX_SURFACES give CRIMSON_EARG.  Results equal
 * crimson_form_boundary_flux's bit for bit.  Every call on one device shares
 * that array: calls that overlap on one device (other threads, other
 * handles' streams) race and must be serialised by the caller. */
#define CRIMSON_FLUX_NOALLOC_MAX_SURFACES 4096
int crimson_form_boundary_flux_noalloc(crimson_form *f, int nflx, double *flxid);

/* Measured ceilings (no counterpart in the reference; measurement only): a
 * plain device copy of copy_bytes (GB/s of read + write) and FP64 / FP32
 * fused multiply-add throughput (TFLOP/s), timed with HIP events. */
int crimson_measure_peaks(size_t copy_bytes, double *copy_gbs, double *fp64_tflops, double *fp32_tflops);

/* Ensembles (no counterpart in the reference, which runs one case per job):
 * the handle's mesh is the union of several instances' meshes (disjoint
 * node ranges), inst(nel) gives each element's 0-based instance and mu(ninst)
 * each instance's viscosity, which replaces crimson_form_params.mu for its
 * elements (and its boundary elements, crimson_form_set_boundary).  ninst =
 * 0 returns to the single viscosity. */
int crimson_form_set_instances(crimson_form *f, const unsigned short *inst,
                               int ninst, const double *mu);

/* The formation into the memLS matrix (crimson_form_interior_to_lhs) captured
 * once into a HIP graph over fixed device arrays, then replayed by
 * crimson_graph_launch (synchronous) for each nonlinear iteration whose
 * inputs were written into the same arrays.  No counterpart in the
 * reference.  No per-phase timing is recorded for replays.  The graph holds
 * the device addresses of the handle's buffers and of lhs's matrix: once a
 * later call has replaced any of them or changed what the capture chose
 * (crimson_form_set_instances with another instance count or 0,
 * crimson_form_set_boundary, crimson_form_set_bc, crimson_form_autotune,
 * crimson_lhs_set_precision, a formation into lhs at another precision, ...),
 * crimson_graph_launch replays nothing and returns CRIMSON_ESTATE; free the
 * graph and capture again.  New viscosities through
 * crimson_form_set_instances with the same instance count keep the graph
 * valid (they are read at each replay).  Free the graph before f and lhs. */
typedef struct crimson_graph crimson_graph;
int crimson_form_graph_create(crimson_form *f, const double *d_y,
                              const double *d_ac, const double *d_x,
                              double *d_Res4, struct crimson_lhs *lhs,
                              crimson_graph **out);
int crimson_graph_launch(crimson_graph *g);
int crimson_graph_free(crimson_graph *g);

/* The NS solve of crimson_ls_solve_dev (LS_TYPE_NS, dof 4) captured into a
 * HIP graph on lhs's matrix and the device array d_Ri (solved in place at
 * each crimson_graph_launch, which then leaves the statistics in ls).  The
 * solve is device-driven: every decision the host takes in memLS's NSSOLVER
 * (memLS.f90:1460-1759: GMRES's Givens rotations and convergence test,
 * CGRAD's iteration count, the least-squares Gram system and its GE, BCPRE)
 * is taken by the GPU, every possible iteration is enqueued and those past
 * convergence return at once; the results equal crimson_ls_solve_dev's.
 * res and incL as in crimson_ls_solve_dev, fixed at capture.  No
 * counterpart in the reference.  The matrix must be held at capture (a later
 * formation into lhs refills the same buffer). */
int crimson_ls_graph_create(struct crimson_lhs *lhs, struct crimson_ls *ls, int dof,
                            double *d_Ri, const int *incL, const double *res,
                            crimson_graph **out);

/* One generalized-alpha timestep (itrdrv.f90:1284-1528, without its host
 * side) captured into a HIP graph: itrPredict (ipred 1), itrBC, then
 * `nonlinear` times itrYAlpha, the formation into lhs's matrix (interior and
 * boundary elements), bc3Res, the ElmpvsQ term of each coupled surface
 * (surface q: nodes d_glob[surf_ptr[q]..surf_ptr[q+1]) 0-based with NABI
 * rows d_nabi(3, .), implicit coefficients c1[q], c2[q], flow into
 * d_q[q]), the device-driven NS solve of crimson_ls_graph_create (faceRes:
 * memLS's face coefficients), itrCorrect and itrBC; then yold, acold <- y,
 * ac.  Each crimson_graph_launch advances one timestep; d_BC may be rewritten
 * between launches (time-dependent Dirichlet values).  For ensembles (the
 * union mesh of crimson_form_set_instances with one coupled surface per
 * instance) the whole timestep of every instance is one launch.  No
 * counterpart in the reference. */
int crimson_step_graph_create(crimson_form *f, struct crimson_lhs *lhs, struct crimson_ls *ls,
                              double *d_yold, double *d_acold, double *d_y, double *d_ac,
                              double *d_yAlpha, double *d_acAlpha, const double *d_x,
                              double *d_Res4, const int *d_iBC, const double *d_BC,
                              int ndofBC, int nsurf, const int *surf_ptr, const int *d_glob,
                              const double *d_nabi, const double *c1, const double *c2,
                              double *d_q, const double *faceRes, int nonlinear,
                              crimson_graph **out);

/* Adjoint of the interior residual (the vector-Jacobian product the
 * estimation subsystem needs for parameter fitting; no counterpart in the
 * reference, which has no derivative code): for a cotangent d_w(nshg,4)
 * laid out like res, d_ybar(nshg,4) = w^T d res/d y with y = (u1,u2,u3,p)
 * and *mubar = w^T d res/d mu, the exact derivative of the residual that
 * crimson_form_interior returns (tau, tauBar, ubar and the reconstructed
 * flux q all differentiated; the acceleration held fixed).  Device arrays,
 * synchronous. */
int crimson_form_vjp_dev(crimson_form *f, const double *d_y, const double *d_ac,
                         const double *d_x, const double *d_w, double *d_ybar,
                         double *mubar);
/* crimson_form_vjp_dev with the acceleration's cotangent as well:
 * d_acbar(nshg,4) = w^T d res/d ac (its fourth column zero; the acceleration
 * enters the momentum residual and the strong residual L, e3res.f90:775-789).
 * The implicit step's adjoint needs it: the tangent of one generalized-alpha
 * step with respect to its acceleration increment is almi d res/d ac +
 * alfi gami Delt d res/d y (itrPC.f90:176-181, 326-334).  No counterpart in
 * the reference. */
int crimson_form_vjp_ac_dev(crimson_form *f, const double *d_y, const double *d_ac,
                            const double *d_x, const double *d_w, double *d_ybar,
                            double *d_acbar, double *mubar);
int crimson_form_free(crimson_form *f);

/* Native genadj (genadj.f90:1-84): node-adjacency CSR including self, rows
 * ascending.  colm must hold nshg+1 ints; rowp must hold maxnnz ints.
 * Returns CRIMSON_EARG if maxnnz is too small (nnz_out holds the need). */
int crimson_genadj(int nshg, int nel, const int *ien, int *colm, int *rowp,
                   int maxnnz, int *nnz_out);

/* ------------------------------------------------------------------ */
/* The steps either side of formation and solve in one nonlinear
 * iteration (itrdrv.f90:1284-1528), on device arrays y(nshg,4),
 * ac(nshg,4) and the memLS increment Res4(4,nshg).  Asynchronous on the
 * formation handle's stream; crimson_form_sync waits for them. */

/* itrPredict, ipred = 1 (itrPC.f90:104-108). */
int crimson_itr_predict_dev(crimson_form *f, const double *d_yold,
                            const double *d_acold, double *d_y, double *d_ac,
                            double gami);
/* itrYAlpha (itrPC.f90:295-365), non-ALE, ndof = 4. */
int crimson_itr_yalpha_dev(crimson_form *f, const double *d_yold,
                           const double *d_acold, const double *d_y,
                           const double *d_ac, double *d_yAlpha,
                           double *d_acAlpha, double alfi, double almi);
/* itrCorrect (itrPC.f90:164-192) with solinc = Res4^T. */
int crimson_itr_correct_dev(crimson_form *f, double *d_y, double *d_ac,
                            const double *d_Res4, double gami, double alfi,
                            double Delt);
/* itrBC (itrbc.f90:1-100, no periodic nodes): velocity codes 1-7 and the
 * pressure bit; d_iBC(nshg), d_BC(nshg,ndofBC) column-major, ndofBC >= 6. */
int crimson_itr_bc_dev(crimson_form *f, const int *d_iBC, const double *d_BC,
                       int ndofBC, double *d_y);
/* Resistance-outlet term of ElmpvsQ (elmgmr.f90:1152-1176): Q = sum NABI.u
 * over the face nodes d_glob (0-based), Res4(1:3,i) += -Q R NABI(i,1:3);
 * d_nabi(3,nface); Q is left in *d_q. */
int crimson_elmpvsq_dev(crimson_form *f, int nface, const int *d_glob,
                        const double *d_nabi, double R, const double *d_y,
                        double *d_Res4, double *d_q);
/* Any coupled surface of ElmpvsQ (elmgmr.f90:1152-1330: resistance,
 * impedance, RCR, netlist, ...) given its 0D model's implicit coefficients
 * for this step, P = c1 Q + c2: Res4(1:3,i) += -(c1 Q + c2) NABI(i,1:3)
 * over d_glob; Q in *d_q.  memLS's face coefficient of the surface is
 * c1 alfi gami Delt (AddElmpvsQFormemLS, :1775-1856). */
int crimson_elmpvsq_coef_dev(crimson_form *f, int nface, const int *d_glob,
                             const double *d_nabi, double c1, double c2,
                             const double *d_y, double *d_Res4, double *d_q);
/* Adjoint of crimson_elmpvsq_coef_dev's residual term for a cotangent
 * d_w4 laid out like Res4(4,nshg): with s = sum over d_glob of w(1:3).NABI
 * and Q = sum NABI.u(d_y), d_ybar(nshg,4) (column-major) gets -c1 s NABI
 * added on the face nodes' velocity components, and d_cbar(2) receives the
 * cotangents (-Q s, -s) of c1 and c2 (of R and of a constant pressure for a
 * resistance).  No counterpart in the reference. */
int crimson_elmpvsq_vjp_dev(crimson_form *f, int nface, const int *d_glob,
                            const double *d_nabi, double c1, const double *d_y,
                            const double *d_w4, double *d_ybar, double *d_cbar);
int crimson_form_sync(crimson_form *f);

/* ------------------------------------------------------------------ */
/* memLS (memLS.f90).                                                   */

enum { CRIMSON_LS_CG = 1, CRIMSON_LS_GMRES = 2, CRIMSON_LS_NS = 3 };
enum { CRIMSON_BC_DIR = 0, CRIMSON_BC_NEU = 1 };


/* memLS_LS_CREATE (memLS.f90:1379-1442).  `present` marks the optional
 * arguments given: 1 relTol, 2 absTol, 4 maxItr, 8 dimKry, 16 relTolIn,
 * 32 absTolIn, 64 maxItrIn. */
int crimson_ls_create(int lsType, int present, double relTol, double absTol,
                      int maxItr, int dimKry, const double *relTolIn,
                      const double *absTolIn, const int *maxItrIn,
                      crimson_ls **out);
int crimson_ls_free(crimson_ls *ls);

/* memLS_LHS_CREATE (memLS.f90:1080-1136), single task: gnNo == nNo,
 * gNodes is the identity (may be NULL), rowPtr(nNo+1) and colPtr(nnz)
 * 1-based (itrdrv.f90:934 passes rowPtr=colm, colPtr=rowp). */
int crimson_lhs_create(int gnNo, int nNo, int nnz, const int *gNodes,
                       const int *rowPtr, const int *colPtr, int nFaces,
                       crimson_lhs **out);
int crimson_lhs_free(crimson_lhs *lhs);

/* memLS_BC_CREATE (memLS.f90:272-316); val(dof,nNo) may be NULL (zeros). */
int crimson_bc_create(crimson_lhs *lhs, int faIn, int nNo, int dof,
                      int bcType, const int *gNodes, const double *val);

/* memLS_SOLVE (memLS.f90:1947-2020) on host arrays: Ri(dof,nNo) in: the
 * right-hand side, out: the solution; Val(dof*dof,nnz) is read (its
 * contents on return are those passed in: the scaling is applied on the
 * GPU copy).  incL(nFaces) and res(nFaces) may be NULL. dof must be 4 for
 * LS_TYPE_NS; 1..4 for GMRES and CG. */
int crimson_ls_solve(crimson_lhs *lhs, crimson_ls *ls, int dof, double *Ri,
                     const double *Val, const int *incL, const double *res);

/* memLS_SOLVE on device data: the matrix is the one already held by `lhs`
 * (from crimson_form_interior_to_lhs or crimson_lhs_set_val), d_Ri(dof,nNo)
 * device in/out. */
int crimson_ls_solve_dev(crimson_lhs *lhs, crimson_ls *ls, int dof,
                         double *d_Ri, const int *incL, const double *res);

/* Uploads Val(dof*dof,nnz) (host, row-major blocks) into the lhs. */
int crimson_lhs_set_val(crimson_lhs *lhs, int dof, const double *Val);

/* Downloads the held matrix as Val(dof*dof,nnz) row-major blocks (the
 * layout solfar.f90:212-235 packs), e.g. after crimson_form_interior_to_lhs. */
int crimson_lhs_get_val(crimson_lhs *lhs, int dof, double *Val);

/* No counterpart in the reference: the mixed-precision form's switch.
 * Converts the held matrix to FP32 (bits = 32) or back to FP64 (bits = 64);
 * with an FP32 matrix every following solve runs its Krylov vectors, SpMV
 * and dot products in FP32 (FP32 accumulation), keeping FP64 for the
 * Hessenberg/Givens least squares, the NS Gram system and the FP64 Ri in
 * and out.  crimson_form_interior_to_lhs with a precision-32 formation
 * handle leaves an FP32 matrix directly. */
int crimson_lhs_set_precision(crimson_lhs *lhs, int bits);
/* The block transpose of src's matrix into dst (both created on the same
 * CSR graph): block (i,j) of dst becomes block (j,i) of src transposed, at
 * src's precision, so that crimson_ls_solve_dev on dst solves with the
 * transposed memLS matrix (the adjoint's transpose solve; memLS has no
 * transpose mode).  dst's faces are its own; the PRECOND scaling, the
 * Dirichlet masks and the rank-one face terms are symmetric.  Waits for
 * src's stream; synchronous. */
int crimson_lhs_transpose_dev(crimson_lhs *dst, const crimson_lhs *src);

/* Stats of the last solve: sub 0 = RI, 1 = GM, 2 = CG; out[0] iNorm,
 * out[1] fNorm, out[2] itr, out[3] suc, out[4] dB, out[5] callD (s),
 * out[6] Resm, out[7] Resc. */
int crimson_ls_stats(const crimson_ls *ls, int sub, double *out8);

/* Measurement hooks for the Krylov kernels on the lhs's stream: one SpMV of
 * the scaled 3x3 momentum block (dof 3) or the full 4x4 matrix (dof 4) on
 * device vectors d_x -> d_y (dof,nNo), repeated `reps` times; returns the
 * average kernel time in ms in *ms (HIP events around the launches). */
int crimson_lhs_spmv_dev(crimson_lhs *lhs, int dof, const double *d_x,
                         double *d_y, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
