// Shared definitions of the formation kernels (form.hip) and their adjoint
// (adjoint.hip): quadrature constants, element/node record layouts, the P1
// metric and e3gijd, and the formation handle.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace {

// 4-point rule (symtet.c:22-23): qp q has barycentric (r,s,t) and
// N_4 = 1 - r - s - t evaluated left to right (uniformP.c:22-25).
constexpr double kA4 = 0.5854101966249685;
constexpr double kB4 = 0.1381966011250150;
constexpr double kQwt = (4.0 / 3.0) * 0.25;  // genint.f90:85
__host__ __device__ constexpr double nq(int q, int a) {
  return a == 3 ? (q == 0 ? 1.0 - kA4 - kB4 - kB4
                          : q == 1 ? 1.0 - kB4 - kA4 - kB4
                          : q == 2 ? 1.0 - kB4 - kB4 - kA4
                                   : 1.0 - kB4 - kB4 - kB4)
                : (a == q ? kA4 : kB4);
}
__host__ __device__ constexpr double sum_n(int a) { return nq(0, a) + nq(1, a) + nq(2, a) + nq(3, a); }
__host__ __device__ constexpr double mm(int a, int b) {
  return nq(0, a) * nq(0, b) + nq(1, a) * nq(1, b) + nq(2, a) * nq(2, b) + nq(3, a) * nq(3, b);
}

// Element record: 32 doubles = two 128-byte lines; node a's S row and
// residual row sit together in half a line at R_A(a).  The shape gradients
// are not stored: the assembly rebuilds them from the element's nodes.
constexpr int kRec = 32;
__host__ __device__ constexpr int R_A(int a) { return 8 * a; }  // S[a][0..3], rl[a][0..3]
constexpr int R_RL = 4;
// Element scalars: 4 doubles per element (Tc, Mu, Cm, tlW), from which with
// the shape gradients every tangent entry is rebuilt.
constexpr int kESc = 4;
// Staged incidence in the FP64 assembly (LDS), and the first 16 floats of
// the FP32 tangent record: g[a][j] at a*3+j (12), tlW, Tc, Mu, Cm; the FP64
// stage then holds node a's S row.
constexpr int R_G = 0;
constexpr int R_TLW = 12, R_TC = 13, R_MU = 14, R_CM = 15;
// Packed node coordinates and velocity, 6 doubles per node (node_pack), the
// gathers of the flux pass and of the assembly's shape gradients.
constexpr int kXU = 6;
// lumped-mass weight of local node a, m_a = sum_q N_a(q) WdetJ (e3q.f90:120-145)
__host__ __device__ inline double flux_weight(double W, int a) {
  return a == 0   ? nq(0, 0) * W + nq(1, 0) * W + nq(2, 0) * W + nq(3, 0) * W
         : a == 1 ? nq(0, 1) * W + nq(1, 1) * W + nq(2, 1) * W + nq(3, 1) * W
         : a == 2 ? nq(0, 2) * W + nq(1, 2) * W + nq(2, 2) * W + nq(3, 2) * W
                  : nq(0, 3) * W + nq(1, 3) * W + nq(2, 3) * W + nq(3, 3) * W;
}
constexpr int kRowLanes = 16;   // lanes per CSR row in the assembly
constexpr int kNodeLanes = 4;   // lanes per node in the flux/residual gathers

// Ensemble instances: per-element instance index and per-instance viscosity
// (both null for a single instance, which uses FormConst::mu).
struct Ens {
  const unsigned short *inst;
  const double *mu;
};
__device__ inline double ens_mu(const Ens &en, int e, double mu) { return en.inst ? en.mu[en.inst[e]] : mu; }

using crimson::xcd_block;

struct FormConst {
  double rho, mu, Delt, lhsFct, lhmFct, Dtgl, dtsfct, taucfct;
};

// e3metric (e3metric.f90:28-81) specialised to P1: dxdxi_ij =
// (x_j,i - x_4,i)/2 since shgl = dN/dxi / 2 (genshp.f90:35-36).
template <class T>
__device__ inline void p1_metric(const T (&xl)[4][3], T (&dx)[3][3], T &WdetJ, T (&g)[4][3]) {
  T d[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) d[i][j] = T(0.5) * xl[j][i] - T(0.5) * xl[3][i];
  dx[0][0] = d[1][1] * d[2][2] - d[2][1] * d[1][2];
  dx[0][1] = d[2][1] * d[0][2] - d[0][1] * d[2][2];
  dx[0][2] = d[0][1] * d[1][2] - d[0][2] * d[1][1];
  const T tmp = T(1) / (dx[0][0] * d[0][0] + dx[0][1] * d[1][0] + dx[0][2] * d[2][0]);
  dx[0][0] *= tmp;
  dx[0][1] *= tmp;
  dx[0][2] *= tmp;
  dx[1][0] = (d[1][2] * d[2][0] - d[1][0] * d[2][2]) * tmp;
  dx[1][1] = (d[0][0] * d[2][2] - d[2][0] * d[0][2]) * tmp;
  dx[1][2] = (d[1][0] * d[0][2] - d[0][0] * d[1][2]) * tmp;
  dx[2][0] = (d[1][0] * d[2][1] - d[1][1] * d[2][0]) * tmp;
  dx[2][1] = (d[2][0] * d[0][1] - d[0][0] * d[2][1]) * tmp;
  dx[2][2] = (d[0][0] * d[1][1] - d[0][1] * d[1][0]) * tmp;
  WdetJ = T(kQwt) / tmp;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    g[0][j] = T(0.5) * dx[0][j];
    g[1][j] = T(0.5) * dx[1][j];
    g[2][j] = T(0.5) * dx[2][j];
    g[3][j] = T(-0.5) * dx[0][j] + T(-0.5) * dx[1][j] + T(-0.5) * dx[2][j];
  }
}

// e3gijd, tet branch (e3stab.f90:493-527)
template <class T>
__device__ inline void gijd_tet(const T (&x)[3][3], T (&gg)[6]) {
  const T c1 = T(1.259921049894873e+00), c2 = T(6.299605249474365e-01);
  T t1, t2, t3;
  t1 = c1 * x[0][0] + c2 * (x[1][0] + x[2][0]);
  t2 = c1 * x[1][0] + c2 * (x[0][0] + x[2][0]);
  t3 = c1 * x[2][0] + c2 * (x[0][0] + x[1][0]);
  gg[0] = x[0][0] * t1 + x[1][0] * t2 + x[2][0] * t3;
  t1 = c1 * x[0][1] + c2 * (x[1][1] + x[2][1]);
  t2 = c1 * x[1][1] + c2 * (x[0][1] + x[2][1]);
  t3 = c1 * x[2][1] + c2 * (x[0][1] + x[1][1]);
  gg[1] = x[0][1] * t1 + x[1][1] * t2 + x[2][1] * t3;
  gg[3] = x[0][0] * t1 + x[1][0] * t2 + x[2][0] * t3;
  t1 = c1 * x[0][2] + c2 * (x[1][2] + x[2][2]);
  t2 = c1 * x[1][2] + c2 * (x[0][2] + x[2][2]);
  t3 = c1 * x[2][2] + c2 * (x[0][2] + x[1][2]);
  gg[2] = x[0][2] * t1 + x[1][2] * t2 + x[2][2] * t3;
  gg[4] = x[0][1] * t1 + x[1][1] * t2 + x[2][1] * t3;
  gg[5] = x[0][0] * t1 + x[1][0] * t2 + x[2][0] * t3;
}

// Node record: 16 doubles = one 128-byte line per node, written by flux_node
// so that the element kernel gathers each node with whole-line loads:
// x(3), u(3), p, ac(3), and the 6 distinct components of q after qpbc
// (q11, q12, q13, q22, q23, q33).
constexpr int kNRec = 16;
constexpr int N_X = 0, N_U = 3, N_P = 6, N_A = 7, N_Q = 10;
// q component k = i + 3j (column-major, the reference's order) -> distinct slot
__host__ __device__ constexpr int qsym(int k) {
  return k == 0 ? 0 : k == 1 ? 1 : k == 2 ? 2 : k == 3 ? 1 : k == 4 ? 3 : k == 5 ? 4 : k == 6 ? 2 : k == 7 ? 4 : 5;
}

inline FormConst make_const(const crimson_form_params &p) {
  FormConst c;
  c.rho = p.rho;
  c.mu = p.mu;
  c.Delt = p.Delt;
  c.lhsFct = p.alfi * p.gami * p.Delt;   // e3lhs.f90:85
  c.lhmFct = p.almi * (1.0 - p.flmpl);   // e3lhs.f90:86
  c.Dtgl = 1.0 / p.Delt;                 // itrPC.f90:71
  c.dtsfct = p.dtsfct;
  c.taucfct = p.taucfct;
  return c;
}

}  // namespace

// The deformable wall's constants (crimson_wall_params and the time
// integration folded as e3bvar uses them).
struct WallConst {
  double h, rhovw;             // SWB(:,1), the wall density
  double d7, d8, d9, d10;      // the isotropic material values SWB(:,7:10)
  double kstiff;               // iwallstiffactor betai Delt^2 alfi (LHSwall's factor besides WdetJb h)
  double mass;                 // almi (1 - flmpl) rhovw h (Wall_LHSfactor's first term)
  double supp, supp_lhs;       // tissSuppStiffCoeff, Delt alfi betai Delt tissSuppStiffCoeff
  double damp, damp_lhs;       // tissSuppDampCoeff, Delt alfi gami tissSuppDampCoeff
  double ring_k, ring_c;       // tissSuppRingStiffCoeff, tissSuppRingDampCoeff
  double ring_klhs, ring_clhs; // betai Delt^2 alfi, Delt alfi gami (the ring tangent's factors)
  int massf;                   // iwallmassfactor
  int isupp, idamp, iring_k, iring_c;
};

// A node with a 1- or 2-component velocity code (bc3lhs): its code, the
// number of elements (interior and boundary) that hold it, and BC(4:6).
struct BclNode {
  int code, count;
  double b[3];
};

// Row schedules of the assembly: the caller's numbering, Morton order of the
// node coordinates, or runs of kSchedTile consecutive rows ordered by the
// Morton code of their centroids.
constexpr int kSchedNumbering = 0, kSchedMorton = 1, kSchedTiles = 2, kSchedTile = 64;

// ---------------------------------------------------------------- handle
struct crimson_form {
  int nshg = 0, numnp = 0, nel = 0, nnz = 0;
  crimson_form_params prm{};
  hipStream_t stream = nullptr;
  int device = 0;  // the device current at setup, which holds every buffer
  crimson::DBuf<int> d_ien, d_inc_ptr, d_inc, d_diagk, d_rowptr, d_ibc;
  crimson::DBuf<int> d_lst_ptr, d_order;  // d_order: assembly row schedule
  crimson::DBuf<int4> d_slot;              // per schedule slot: the row's offsets (slot_table_kernel)
  int sched = 1;                            // assembly row schedule (kSched*; set at setup)
  int asm_chunk = 16;                       // incidences per staged chunk of the FP64 assembly
  unsigned asm_xcd = 0;                     // assembly workgroups dealt per XCD turn (0: one range each)
  crimson::DBuf<unsigned short> d_lst;
  crimson::DBuf<int4> d_ienp;    // per incidence (node-major): the element's four nodes in local order
  crimson::DBuf<double> d_xu;    // (kXU, nshg): x and u of each node (node_pack)
  crimson::DBuf<double> d_node, d_q, d_rec, d_esc;  // node records; q; element records and scalars
  crimson::DBuf<float> d_rec32;  // FP32 tangent records (precision 32)
  crimson::DBuf<double> s_y, s_ac, s_x, s_res, s_K, s_P, s_q;  // staging of the host API
  bool has_ibc = false;
  // partial velocity BCs (crimson_form_set_bc): bc3lhs on the tangent after
  // the assembly, bc3res on the residual; host copies kept when npartial > 0
  int npartial = 0, ndofbc = 0;
  std::vector<int> h_ibc, h_rp, h_cols, h_ninc, h_nbnd;
  std::vector<double> h_bc;
  crimson::DBuf<double> d_bc;      // BC(nshg, ndofbc)
  crimson::DBuf<int4> d_bcl;       // blocks touched: k, row's / column's partial index (-1: none), diagonal
  crimson::DBuf<BclNode> d_bcn;    // the partial nodes
  int nbcl = 0;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  float ms[4] = {0, 0, 0, 0};
  // adjoint (VJP) workspaces, allocated on first use
  crimson::DBuf<double> d_vrec, d_vrec2, d_qbar, d_invm, d_mupart;
  // ensemble instances (crimson_form_set_instances)
  crimson::DBuf<unsigned short> d_inst;
  crimson::DBuf<double> d_mu_inst;
  int ninst = 0;
  Ens ens() const { return Ens{ninst ? d_inst.p : nullptr, ninst ? d_mu_inst.p : nullptr}; }
  crimson::DBuf<unsigned short> d_binst;  // instance of each boundary element (form_boundary_instances)
  Ens bens() const { return Ens{ninst && nelb ? d_binst.p : nullptr, ninst ? d_mu_inst.p : nullptr}; }
  // boundary elements (crimson_form_set_boundary)
  int nelb = 0, nbnode = 0, nqb = 3, nsrf = -1;
  double influx = 0.5;
  crimson::DBuf<int> d_ienb;      // (4, nelb) 0-based, face nodes first
  crimson::DBuf<int> d_bcode;     // (2, nelb): iBCB code bits, srfID
  crimson::DBuf<double> d_bcb;    // BCB(nelb, 3, 5): mass flux, pressure, traction (zero past ndBCB)
  crimson::DBuf<int> d_permit;    // flow-permitted flag by srfID, 0..nsrf
  crimson::DBuf<double> d_brec;   // (12, nelb): rl of the three face nodes
  crimson::DBuf<double> d_bflx;   // (5, nelb): surface integrals of the last formation (flxID terms)
  crimson::DBuf<int> d_sptr, d_selem;  // boundary elements by |srfID|, 0..maxsrf
  crimson::DBuf<double> d_flxout;
  int maxsrf = 0;
  crimson::DBuf<int> d_bnode, d_binc_ptr, d_binc;  // boundary nodes; their incidences e << 2 | a
  int nvnode = 0;                                  // the adjoint's: all four nodes of each element
  crimson::DBuf<int> d_vnode, d_vptr, d_vinc;
  crimson::DBuf<double> d_bvrec;                   // boundary VJP records
  // deformable walls (crimson_form_set_wall)
  bool wall = false;
  unsigned long long wall_gen = 0;  // bumped by every crimson_form_set_wall (captured graphs check it)
  WallConst wc{};
  int nwf = 0, nwn = 0, nwk = 0;   // wall faces, wall nodes, matrix blocks the wall tangent touches
  crimson::DBuf<int> d_wface;      // boundary-element index of each wall face
  crimson::DBuf<int> d_wnode;      // the wall nodes (mWNodes), 0-based
  crimson::DBuf<int> d_wtag;       // (3, nwf): the face nodes' ring bits
  crimson::DBuf<double> d_uref;    // uref(nshg,3)
  const double *d_u = nullptr;     // the displacement of the next formations (the caller's)
  double *d_wuold = nullptr, *d_wu = nullptr;  // u at t_n and t_{n+1} a captured timestep updates
  crimson::DBuf<double> d_wres;    // (9, nwf): the wall rows of the three face nodes
  crimson::DBuf<double> d_wrec;    // (81, nwf): xKebe(9, a, b) of the face node pairs
  crimson::DBuf<int> d_wk, d_wptr, d_wsrc;  // the tangent gather: block k, its (face, pair) sources
  std::vector<int> h_wrp, h_wcols;          // the CSR graph (host), kept for the wall's set-up
};

namespace crimson {
// (Re)builds the bc3lhs block table from the handle's host copies and BC;
// a no-op until BC is set.
int form_bc3lhs_table(crimson_form *f);
int form_flux_pass(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x, double *d_q,
                   double *d_invm);
// Boundary elements' residual rows added to res (mode 0: res(nshg,4); mode
// 1: Res4(4,nshg)) from the node records of the current formation; a no-op
// without boundary elements.  Enqueued on the handle's stream.
int form_boundary_pass(crimson_form *f, double *d_res, int mode);
// The wall faces' rows and, when lhs, their tangent added into the matrix
// (mode 0: lhsK(9,nnz); mode 1: the FP64 or FP32 planes of o), from the
// node records of the current formation; enqueued before bc3lhs.  A no-op
// without a wall.
int form_wall_pass(crimson_form *f, int lhs, int mode, double *lhsK, double *V, float *V32, size_t nnz);
// brec += the wall rows (after bnd_elem wrote brec).
int form_wall_rows(crimson_form *f);
// The ensemble instance of each boundary element (after either of
// crimson_form_set_boundary / crimson_form_set_instances changes).
int form_boundary_instances(crimson_form *f);
// Adjoint of the boundary rows: ybar(nshg,4) += w^T d rl_bnd/d y from the
// node records of the last flux pass; *mubar = w^T d rl_bnd/d mu.  d_bvrec:
// a workspace of form_boundary_vjp_doubles(f) doubles, or null for the
// handle's own (allocated on first use).
int form_boundary_vjp(crimson_form *f, const double *d_w, double *d_ybar, double *mubar, double *d_bvrec);
size_t form_boundary_vjp_doubles(const crimson_form *f);
// Device workspaces of one VJP (crimson_form_vjp_dev): the element records,
// qbar, 1/rmass, the mu partial sums and the boundary records (null without
// boundary elements), sized by vjp_work_sizes in doubles.
struct VjpWork {
  double *vrec, *vrec2, *qbar, *invm, *mupart, *bvrec;
};
void vjp_work_sizes(const crimson_form *f, size_t n[6]);
int vjp_run(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x, const double *d_w,
            double *d_ybar, double *d_acbar, double *mubar, const VjpWork &w);
}  // namespace crimson
