// memLS on gfx950: PRECOND, NSSOLVER (GMRES + CG on the Schur complement +
// least-squares combination), GMRESV and CGRADV/CGRADS (memLS.f90).
//
// Design:
//  * The matrix lives as vdof*vdof planes of nnz doubles (plane m = entry
//    (m / vdof, m % vdof) of every row-major block).  Each plane is read with
//    16 lanes per CSR row, so every load instruction covers 128 contiguous
//    bytes.
//  * PRECOND's symmetric diagonal scaling is not written back into Val: every
//    SpMV applies y_r(i) = W_r(i) sum_j A_rc(i,j) W_c(j) x_c(j) + fix_r(i) x_r(i)
//    where fix marks the diagonal entries memLS forces to 1 (zero diagonal,
//    Dirichlet-masked DOF; memLS.f90:1815-1820,1923-1929).  This saves the two
//    read+write sweeps over Val per solve.
//  * Dots and norms are two-stage deterministic reductions (per-workgroup
//    partials in a fixed order, then one workgroup per result).
//  * GMRES orthogonalises with classical Gram-Schmidt and a DGKS second pass
//    (taken when ||w'|| < ||w||/sqrt 2), computing all i projections of a
//    step in one sweep.  memLS uses modified Gram-Schmidt (memLS.f90:743-746);
//    both produce the same Krylov space, and solutions agree to the solver
//    tolerance.  Givens rotations, the convergence test and the triangular
//    solve run on the host in FP64 exactly as memLS.f90:750-775.
//  * CG iterations run device-resident with a device-side convergence flag;
//    the host enqueues them in batches and reads the flag between batches.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "common.hpp"

namespace {

constexpr int kRowLanes = 16;  // lanes per CSR row in the SpMV kernels
constexpr int kRedThreads = 256;
constexpr int kChunk = 1024;   // vector entries per workgroup in the dot sweeps
// Row kernels deal their workgroups to the XCDs in turns of this many
// (crimson::xcd_block(chunk)): rows gather the vector at neighbours a mesh
// layer away, which the shared MALL then holds for all XCDs.  One FP64
// timestep of the 64x64x320 cylinder: 2036 ms with one contiguous range per
// XCD, 1966 / 1918 / 1916 / 1915 / 1929 ms with turns of 4 / 8 / 16 / 32 /
// 64 for the Schur CG's two row kernels, 1912 ms with the SpMV too.
constexpr unsigned kRowXcdChunk = 16;

struct SubLs {
  double absTol = 0, relTol = 0, iNorm = 0, fNorm = 0, dB = 0, callD = 0;
  int suc = 0, mItr = 0, sD = 0, itr = 0;
};

// A scaled view of a sub-block (rows r0.., cols c0..) of the stored matrix.
// T is the arithmetic type of the solve: the matrix planes, W, the Krylov
// vectors and every accumulation (FP64, or FP32 in the mixed form).
template <class T>
struct MatView {
  const T *base;  // plane 0
  size_t nnz;
  int plane[16];  // plane index of (r, c), r < NR, c < NC: plane[r * 4 + c]
  const T *W;     // W(wdof, nNo)
  const T *fix;   // fix(wdof, nNo) or null
  int wdof, r0, c0;
};

// The column of block k of `row`: stored as is (int) or as its offset from
// the row (short, lhs->d_col16).
__device__ inline int col_at(const int *__restrict__ col, int k, int row) { return col[k]; }
__device__ inline int col_at(const short *__restrict__ col, int k, int row) { return row + (int)col[k]; }
// One 4-wide node entry (32 bytes in FP64, 16 in FP32; aligned) in 16-byte
// pieces.
__device__ inline void load4(const double *__restrict__ p, double (&z)[4]) {
  const double2 a = reinterpret_cast<const double2 *>(p)[0], b = reinterpret_cast<const double2 *>(p)[1];
  z[0] = a.x;
  z[1] = a.y;
  z[2] = b.x;
  z[3] = b.y;
}
__device__ inline void load4(const float *__restrict__ p, float (&z)[4]) {
  const float4 a = *reinterpret_cast<const float4 *>(p);
  z[0] = a.x;
  z[1] = a.y;
  z[2] = a.z;
  z[3] = a.w;
}
__device__ inline void store4(double *__restrict__ p, const double (&z)[4]) {
  reinterpret_cast<double2 *>(p)[0] = make_double2(z[0], z[1]);
  reinterpret_cast<double2 *>(p)[1] = make_double2(z[2], z[3]);
}
__device__ inline void store4(float *__restrict__ p, const float (&z)[4]) {
  *reinterpret_cast<float4 *>(p) = make_float4(z[0], z[1], z[2], z[3]);
}

// LANES lanes per CSR row: 16 for the 3x3 and 4x4 products, 4 for the
// narrow ones (G, D, L) whose per-entry work is too small to feed 16 lanes.
template <int NR, int NC, class T, int LANES, class IX>
__global__ void __launch_bounds__(256) spmv_kernel(int nNo, const int *__restrict__ rowptr,
                                                   const IX *__restrict__ col, MatView<T> A,
                                                   const T *__restrict__ x, T *__restrict__ y, T ya,
                                                   const T *__restrict__ z, T za,
                                                   const double *__restrict__ skip) {
  if (skip && *skip != 0.0) return;
  const int gid = crimson::xcd_block(kRowXcdChunk) * blockDim.x + threadIdx.x;
  const int row = gid / LANES;
  const int lane = threadIdx.x % LANES;
  T acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = T(0);
  if (row < nNo) {
    const int ke = rowptr[row + 1];
    for (int k = rowptr[row] + lane; k < ke; k += LANES) {
      const int j = col_at(col, k, row);
      T xv[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) xv[c] = A.W[A.c0 + c + (size_t)A.wdof * j] * x[c + (size_t)NC * j];
#pragma unroll
      for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[r] += A.base[(size_t)A.plane[r * 4 + c] * A.nnz + k] * xv[c];
    }
  }
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1)
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] += __shfl_xor(acc[r], off, LANES);
  if (row < nNo && lane == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      T v = A.W[A.r0 + r + (size_t)A.wdof * row] * acc[r];
      if (A.fix && NR == NC && A.r0 == A.c0) v += A.fix[A.r0 + r + (size_t)A.wdof * row] * x[r + (size_t)NC * row];
      v *= ya;
      if (z) v += za * z[r + (size_t)NR * row];
      y[r + (size_t)NR * row] = v;
    }
  }
}

// The Schur product of NSSOLVER's CG (memLS.f90:1719-1735) after G P:
// SP = L P - Gt t1 with t1 = G P, both scaled views of the same CSR rows, in
// one pass: SP_i = W_i sum_j L_ij (W_j P_j) + fix_i P_i - W_i sum_j Gt_ij (W_j t1_j)
// with Ps = W o P and W o t1 precomputed, gathered per column as one entry
// Z(0:3, j) = (Ps_j, W o t1_j) (the same operations, in the same order, as
// the two separate products).
constexpr int kNarrowLanes = 4;

// A coupled face of the Schur CG (ADDBCMUL's PRE, memLS.f90:1719-1735):
// its nodes, valM, coefficient -res/(1 + res nS), and the face dot
// S_f = valM_f . t1 restated as a functional of the G product's operand,
// S_f = sum_j q_j Ps_j over the support sup (the columns of the face nodes'
// rows), q_j = sum over the pairs (a, k) of column j of
// valM(i, a) W_u(i, glob(a)) G_i(k), i < min(dof, 3).
struct FaceDev {
  const int *glob;
  const double *valM;
  double coef;
  const double *coefp;  // the coefficient on the device (device-driven solves), or null: coef
  double res;           // the face's coupling coefficient (ADDBCMUL without PRE)
  int nNo, dof;
  const int *sup;
  const double *q;
  int nsup;
};
// Workgroups per face computing S_f's partial sums beside the G product.
constexpr int kFaceDotBlocks = 4;

// The Schur operator's vectors pre-multiplied by their column scaling, so
// that its two row kernels gather no W: Ps = W_p o P (written by the CG's
// direction update, scale_p_kernel for the first direction); the G product
// writes t1 = W_u o (G Ps) and, per node i, Z(0:3, i) = (Ps_i, W_u o t1_i):
// the Schur pass gathers both of its operands of a column as one 4-wide
// entry (two 16-byte loads in FP64, one in FP32).  Same products in the
// same order as the unscaled kernels.

// Wave-level sum (64 lanes) then across the 4 waves of a 256-thread group.
template <class T>
__device__ inline T block_sum(T v, T *sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  T t = T(0);
  if (threadIdx.x == 0) t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return t;
}

// block_sum's value in every thread of the workgroup.
template <class T>
__device__ inline T block_sum_all(T v, T *sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  const T t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return t;
}

// Entries per lane loaded ahead in the narrow row kernels: the column
// indices of kNarrowUnroll entries (k, k + 4, ...) are loaded together, then
// the gathered vector entries and matrix values, so a lane waits for one
// index load and one gather per kNarrowUnroll entries instead of per entry.
// Indices past the row end are clamped to its last entry (in bounds; the
// product is not added), so the accumulation order is the plain loop's.
constexpr int kNarrowUnroll = 4;

// `done` (may be null): the CG's converged flag; the remaining iterations of
// a batch launched after convergence return at once.
template <class T>
__global__ void scale_p_kernel(int n, const T *__restrict__ W, int wdof, int off, const T *__restrict__ x,
                               T *__restrict__ y, const double *__restrict__ done) {
  if (done && *done != 0.0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = W[off + (size_t)wdof * i] * x[i];
}
// The first nfc * kFaceDotBlocks workgroups form the coupled faces' dots
// from Ps (fpart[f * kFaceDotBlocks + b]: the b-th workgroup's share, entries
// s = b * 256 + t + 1024 u), so that no launch between the two row kernels
// waits for t1; the rest form t1.
template <class T, class IX, int U = kNarrowUnroll>
__global__ void __launch_bounds__(256) g_product_kernel(int nNo, const int *__restrict__ rowptr,
                                                        const IX *__restrict__ col, MatView<T> G,
                                                        const T *__restrict__ Ps, T *__restrict__ t1,
                                                        T *__restrict__ t1s, int nfc,
                                                        const FaceDev *__restrict__ fd,
                                                        double *__restrict__ fpart,
                                                        const double *__restrict__ done) {
  if (done && *done != 0.0) return;
  const unsigned nx = (unsigned)nfc * kFaceDotBlocks;
  if (blockIdx.x < nx) {
    __shared__ double sh[4];
    const int f = blockIdx.x / kFaceDotBlocks, b = blockIdx.x % kFaceDotBlocks;
    const int *sup = fd[f].sup;
    const double *q = fd[f].q;
    const int ns = fd[f].nsup;
    constexpr int kStride = 256 * kFaceDotBlocks;
    double s = 0.0;
    for (int s0 = b * 256 + threadIdx.x; s0 < ns; s0 += kStride * 4) {
      int j[4];
      double qv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = min(s0 + kStride * u, ns - 1);
        j[u] = sup[e];
        qv[u] = q[e];
      }
      T pv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) pv[u] = Ps[j[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (s0 + kStride * u < ns) s += qv[u] * (double)pv[u];
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) fpart[blockIdx.x] = s;
    return;
  }
  const int gid = crimson::xcd_block(kRowXcdChunk, blockIdx.x - nx, gridDim.x - nx) * blockDim.x + threadIdx.x;
  const int row = gid / kNarrowLanes;
  const int lane = threadIdx.x % kNarrowLanes;
  T acc[3] = {T(0), T(0), T(0)};
  if (row < nNo) {
    const int ke = rowptr[row + 1];
    const T *g0 = G.base + (size_t)G.plane[0] * G.nnz, *g1 = G.base + (size_t)G.plane[4] * G.nnz,
            *g2 = G.base + (size_t)G.plane[8] * G.nnz;
    for (int k = rowptr[row] + lane; k < ke; k += kNarrowLanes * U) {
      int j[U], kc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        kc[u] = min(k + kNarrowLanes * u, ke - 1);
        j[u] = col_at(col, kc[u], row);
      }
      T xv[U], a0[U], a1[U], a2[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        xv[u] = Ps[j[u]];
        a0[u] = g0[kc[u]];
        a1[u] = g1[kc[u]];
        a2[u] = g2[kc[u]];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (k + kNarrowLanes * u < ke) {
          acc[0] += a0[u] * xv[u];
          acc[1] += a1[u] * xv[u];
          acc[2] += a2[u] * xv[u];
        }
    }
  }
#pragma unroll
  for (int off = kNarrowLanes / 2; off > 0; off >>= 1)
#pragma unroll
    for (int r = 0; r < 3; ++r) acc[r] += __shfl_xor(acc[r], off, kNarrowLanes);
  if (row < nNo && lane == 0) {
    T z[4];
    z[0] = Ps[row];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const T w = G.W[G.r0 + r + (size_t)G.wdof * row];
      T v = w * acc[r];
      v *= T(1);
      t1[r + (size_t)3 * row] = v;
      z[1 + r] = w * v;
    }
    store4(t1s + (size_t)4 * row, z);
  }
}
template <class T, class IX, int U = kNarrowUnroll>
__global__ void __launch_bounds__(256) schur_kernel(int nNo, const int *__restrict__ rowptr,
                                                    const IX *__restrict__ col, MatView<T> Lv, MatView<T> Gv,
                                                    const T *__restrict__ P, const T *__restrict__ Ps,
                                                    const T *__restrict__ t1s, T *__restrict__ SP,
                                                    const double *__restrict__ done) {
  if (done && *done != 0.0) return;
  const int gid = crimson::xcd_block(kRowXcdChunk) * blockDim.x + threadIdx.x;
  const int row = gid / kNarrowLanes;
  const int lane = threadIdx.x % kNarrowLanes;
  T aL = T(0), aG = T(0);
  if (row < nNo) {
    const int ke = rowptr[row + 1];
    const T *lp = Lv.base + (size_t)Lv.plane[0] * Lv.nnz;
    const T *d0 = Gv.base + (size_t)Gv.plane[0] * Gv.nnz, *d1 = Gv.base + (size_t)Gv.plane[1] * Gv.nnz,
            *d2 = Gv.base + (size_t)Gv.plane[2] * Gv.nnz;
    for (int k = rowptr[row] + lane; k < ke; k += kNarrowLanes * U) {
      int j[U], kc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        kc[u] = min(k + kNarrowLanes * u, ke - 1);
        j[u] = col_at(col, kc[u], row);
      }
      T l[U], ps[U], b0[U], b1[U], b2[U], x0[U], x1[U], x2[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        l[u] = lp[kc[u]];
        b0[u] = d0[kc[u]];
        b1[u] = d1[kc[u]];
        b2[u] = d2[kc[u]];
        T z[4];
        load4(t1s + (size_t)4 * j[u], z);  // (Ps_j, W_u o t1_j)
        ps[u] = z[0];
        x0[u] = z[1];
        x1[u] = z[2];
        x2[u] = z[3];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (k + kNarrowLanes * u < ke) {
          aL += l[u] * ps[u];
          aG += b0[u] * x0[u];
          aG += b1[u] * x1[u];
          aG += b2[u] * x2[u];
        }
    }
  }
#pragma unroll
  for (int off = kNarrowLanes / 2; off > 0; off >>= 1) {
    aL += __shfl_xor(aL, off, kNarrowLanes);
    aG += __shfl_xor(aG, off, kNarrowLanes);
  }
  if (row < nNo && lane == 0) {
    T v = Lv.W[Lv.r0 + (size_t)Lv.wdof * row] * aL;
    if (Lv.fix) v += Lv.fix[Lv.r0 + (size_t)Lv.wdof * row] * P[row];
    v *= T(1);
    const T t2 = Gv.W[Gv.r0 + (size_t)Gv.wdof * row] * aG;
    SP[row] = v + T(-1) * t2;
  }
}


// partial[blk][j] = sum over this block's chunk of U_j . v for j < m, and
// partial[blk][m] = v . v when self != 0.  U_j = U0 + j * ustride.  Vectors
// of type T, products and sums in Acc.
template <class T, class Acc>
__global__ void __launch_bounds__(256) mdot_kernel(size_t L, const T *__restrict__ v, const T *__restrict__ U0,
                                                   size_t ustride, int m, int self, Acc *__restrict__ partial,
                                                   int pstride, const int *__restrict__ gate,
                                                   const double *__restrict__ skip) {
  if ((gate && *gate == 0) || (skip && *skip != 0.0)) return;
  __shared__ Acc sh[4];
  const size_t base = (size_t)blockIdx.x * kChunk;
  Acc vv[kChunk / 256];
#pragma unroll
  for (int t = 0; t < kChunk / 256; ++t) {
    const size_t e = base + threadIdx.x + 256 * t;
    vv[t] = e < L ? (Acc)v[e] : Acc(0);
  }
  Acc *out = partial + (size_t)blockIdx.x * pstride;
  for (int j = 0; j < m; ++j) {
    const T *u = U0 + ustride * j;
    Acc s = Acc(0);
#pragma unroll
    for (int t = 0; t < kChunk / 256; ++t) {
      const size_t e = base + threadIdx.x + 256 * t;
      if (e < L) s += (Acc)u[e] * vv[t];
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[j] = s;
  }
  if (self) {
    Acc s = Acc(0);
#pragma unroll
    for (int t = 0; t < kChunk / 256; ++t) s += vv[t] * vv[t];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[m] = s;
  }
}

// out[j] = sum_blk partial[blk][j], j < m, one workgroup per j, fixed order
// (summed in T, stored as double for the host-side FP64 scalars).
template <class T>
__global__ void __launch_bounds__(256) reduce_kernel(int nblk, const T *__restrict__ partial, int pstride,
                                                     double *__restrict__ out, const int *__restrict__ gate,
                                                     const double *__restrict__ skip) {
  __shared__ T sh[4];
  const int j = blockIdx.x;
  if (skip && *skip != 0.0) return;
  if (gate && *gate == 0) {
    if (threadIdx.x == 0) out[j] = 0.0;
    return;
  }
  T s = T(0);
  for (int b = threadIdx.x; b < nblk; b += 256) s += partial[(size_t)b * pstride + j];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[j] = (double)s;
}

// y -= sum_j coef[j] U_j (coef on the device), then the block's partial of
// y . y into partial[blk * pstride].
template <class T>
__global__ void __launch_bounds__(256) msub_norm_kernel(size_t L, T *__restrict__ y, const T *__restrict__ U0,
                                                        size_t ustride, int m, const double *__restrict__ coef,
                                                        T *__restrict__ partial, int pstride,
                                                        const int *__restrict__ gate,
                                                        const double *__restrict__ skip) {
  if ((gate && *gate == 0) || (skip && *skip != 0.0)) return;
  __shared__ T sh[4];
  __shared__ T cs[256];
  for (int j = threadIdx.x; j < m && j < 256; j += 256) cs[j] = (T)coef[j];
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kChunk;
  T s = T(0);
#pragma unroll
  for (int t = 0; t < kChunk / 256; ++t) {
    const size_t e = base + threadIdx.x + 256 * t;
    if (e < L) {
      T w = y[e];
      for (int j = 0; j < m; ++j) w -= cs[j] * U0[ustride * j + e];
      y[e] = w;
      s += w * w;
    }
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[(size_t)blockIdx.x * pstride] = s;
}

// y += sum_j coef[j] U_j, coef on the device.
template <class T>
__global__ void maxpy_kernel(size_t L, T *__restrict__ y, const T *__restrict__ U0, size_t ustride, int m,
                             const double *__restrict__ coef, double sign) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L) return;
  T w = y[e];
  for (int j = 0; j < m; ++j) w += (T)(sign * coef[j]) * U0[ustride * j + e];
  y[e] = w;
}

// y = a x + b y
template <class T>
__global__ void axpby_kernel(size_t L, T a, const T *__restrict__ x, T b, T *__restrict__ y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L) return;
  y[e] = a * x[e] + (b == T(0) ? T(0) : b * y[e]);
}

// y = x * w elementwise (types may differ: the FP64 right-hand side in and
// the solution out of an FP32 solve)
template <class TX, class TW, class TY>
__global__ void mul_kernel(size_t L, const TX *__restrict__ x, const TW *__restrict__ w, TY *__restrict__ y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L) return;
  y[e] = (TY)(x[e] * w[e]);
}

// Strided component extraction/insert between (dof,N) and (n,N) vectors.
template <class T>
__global__ void extract_kernel(int N, int dof, int c0, int nc, const T *__restrict__ src, T *__restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int c = 0; c < nc; ++c) dst[c + (size_t)nc * i] = src[c0 + c + (size_t)dof * i];
}
template <class T>
__global__ void insert_kernel(int N, int dof, int c0, int nc, const T *__restrict__ src, T *__restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int c = 0; c < nc; ++c) dst[c0 + c + (size_t)dof * i] = src[c + (size_t)nc * i];
}

// PRECOND part 1: W from the diagonal, fix for zero diagonals.
template <class T>
__global__ void precond_diag_kernel(int N, int dof, int vdof, const T *__restrict__ V, size_t nnz,
                                    const int *__restrict__ diag, T *__restrict__ W, T *__restrict__ fix) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int d = diag[i];
  for (int r = 0; r < dof; ++r) {
    const T a = V[(size_t)(r * vdof + r) * nnz + d];
    W[r + (size_t)dof * i] = (a == T(0)) ? T(1) : T(1) / sqrt(fabs(a));
    fix[r + (size_t)dof * i] = (a == T(0)) ? T(1) : T(0);
  }
}
// Dirichlet masking of W by a face's val (memLS.f90:1831-1841)
template <class T>
__global__ void precond_face_kernel(int n, int dof, int fdof, const int *__restrict__ glob,
                                    const double *__restrict__ val, T *__restrict__ W) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int nc = fdof < dof ? fdof : dof;
  const int Ac = glob[a];
  for (int i = 0; i < nc; ++i) W[i + (size_t)dof * Ac] *= (T)val[i + (size_t)fdof * a];
}
// After masking: diagonal entries that scale to zero are forced to one.
template <class T>
__global__ void precond_fix_kernel(int N, int dof, int vdof, const T *__restrict__ V, size_t nnz,
                                   const int *__restrict__ diag, const T *__restrict__ W, T *__restrict__ fix) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int d = diag[i];
  for (int r = 0; r < dof; ++r) {
    const T a = V[(size_t)(r * vdof + r) * nnz + d];
    const T w = W[r + (size_t)dof * i];
    if (a == T(0) || (a * w) * w == T(0)) fix[r + (size_t)dof * i] = T(1);
  }
}
// valM = val * W for a coupled face (kept in FP64: faces are small)
template <class T>
__global__ void face_valm_kernel(int n, int dof, int fdof, const int *__restrict__ glob,
                                 const double *__restrict__ val, const T *__restrict__ W,
                                 double *__restrict__ valM) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int nc = fdof < dof ? fdof : dof;
  for (int i = 0; i < nc; ++i)
    valM[i + (size_t)fdof * a] = val[i + (size_t)fdof * a] * (double)W[i + (size_t)dof * glob[a]];
}
// S = sum_a sum_i valM(i,a) X(i, glob(a)) in one 1024-thread workgroup
// (faces are small): per-wave sums, then the 16 wave sums in a fixed order.
template <class T>
__global__ void __launch_bounds__(1024) face_dot_kernel(int n, int xdof, int fdof, const int *__restrict__ glob,
                                                        const double *__restrict__ valM, const T *__restrict__ X,
                                                        double *__restrict__ out, const double *__restrict__ skip) {
  if (skip && *skip != 0.0) return;
  __shared__ double sh[16];
  const int nc = fdof < xdof ? fdof : xdof;
  double s = 0.0;
  for (int a = threadIdx.x; a < n; a += 1024)
    for (int i = 0; i < nc; ++i) s += valM[i + (size_t)fdof * a] * (double)X[i + (size_t)xdof * glob[a]];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += sh[w];
    *out = t;
  }
}
// Y(i, glob(a)) += valM(i,a) * coef * S (coef from *coefp when given)
template <class T>
__global__ void face_axpy_kernel(int n, int xdof, int fdof, const int *__restrict__ glob,
                                 const double *__restrict__ valM, const double *__restrict__ S, double coef,
                                 T *__restrict__ Y, const double *__restrict__ coefp,
                                 const double *__restrict__ skip) {
  if (skip && *skip != 0.0) return;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int nc = fdof < xdof ? fdof : xdof;
  const double s = (coefp ? *coefp : coef) * *S;
  for (int i = 0; i < nc; ++i) Y[i + (size_t)xdof * glob[a]] += (T)(valM[i + (size_t)fdof * a] * s);
}
// A coupled face on the device, with its PRE coefficient -res/(1 + res nS)
// (memLS.f90:1096-1118 with BCPRE's nS).
// ADDBCMUL's PRE on the Schur CG's G product t1 (3, N): face by face
// t1 += coef_f S_f valM_f, with Z(1:3) = W_u o t1 rewritten at the same nodes
// (one launch per face, since faces may share nodes).  S_f, every face's dot
// of the unmodified t1, is the sum of its kFaceDotBlocks partials from the G
// product (summed here in a fixed order).
template <class T>
__global__ void __launch_bounds__(256) face_pre_apply_kernel(FaceDev F, const double *__restrict__ part,
                                                             T *__restrict__ t1, T *__restrict__ t1s,
                                                             const T *__restrict__ W, int wdof,
                                                             const double *__restrict__ done) {
  if (done && *done != 0.0) return;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= F.nNo) return;
  const int nc = F.dof < 3 ? F.dof : 3;
  double S = 0.0;
#pragma unroll
  for (int b = 0; b < kFaceDotBlocks; ++b) S += part[b];
  const double sc = (F.coefp ? *F.coefp : F.coef) * S;
  const size_t g = F.glob[a];
  for (int i = 0; i < nc; ++i) t1[i + 3 * g] += (T)(F.valM[i + (size_t)F.dof * a] * sc);
#pragma unroll
  for (int c = 0; c < 3; ++c) t1s[1 + c + 4 * g] = W[c + (size_t)wdof * g] * t1[c + 3 * g];
}
// The same three face operations for every coupled face in one launch, when
// the faces share no node (then the faces' updates commute): one
// 1024-thread workgroup per face for the dots (face_dot_kernel's order), and
// a grid row per face for the updates.  Per face the arithmetic is the
// single-face kernels'.
template <class T>
__global__ void __launch_bounds__(1024) face_dot_multi_kernel(const FaceDev *__restrict__ fd, int xdof,
                                                              const T *__restrict__ X, double *__restrict__ out,
                                                              const double *__restrict__ skip) {
  if (skip && *skip != 0.0) return;
  __shared__ double sh[16];
  const FaceDev F = fd[blockIdx.x];
  const int nc = F.dof < xdof ? F.dof : xdof;
  double s = 0.0;
  for (int a = threadIdx.x; a < F.nNo; a += 1024)
    for (int i = 0; i < nc; ++i) s += F.valM[i + (size_t)F.dof * a] * (double)X[i + (size_t)xdof * F.glob[a]];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += sh[w];
    out[blockIdx.x] = t;
  }
}
template <class T>
__global__ void face_axpy_multi_kernel(const FaceDev *__restrict__ fd, int xdof, const double *__restrict__ S,
                                       int pre, T *__restrict__ Y, const double *__restrict__ skip) {
  if (skip && *skip != 0.0) return;
  const FaceDev F = fd[blockIdx.y];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= F.nNo) return;
  const int nc = F.dof < xdof ? F.dof : xdof;
  const double s = (pre ? (F.coefp ? *F.coefp : F.coef) : F.res) * S[blockIdx.y];
  for (int i = 0; i < nc; ++i) Y[i + (size_t)xdof * F.glob[a]] += (T)(F.valM[i + (size_t)F.dof * a] * s);
}
template <class T>
__global__ void __launch_bounds__(256) face_pre_apply_multi_kernel(const FaceDev *__restrict__ fd,
                                                                   const double *__restrict__ part,
                                                                   T *__restrict__ t1, T *__restrict__ t1s,
                                                                   const T *__restrict__ W, int wdof,
                                                                   const double *__restrict__ done) {
  if (done && *done != 0.0) return;
  const FaceDev F = fd[blockIdx.y];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= F.nNo) return;
  const int nc = F.dof < 3 ? F.dof : 3;
  const double *pf = part + (size_t)kFaceDotBlocks * blockIdx.y;
  double Sv = 0.0;
#pragma unroll
  for (int b = 0; b < kFaceDotBlocks; ++b) Sv += pf[b];
  const double sc = (F.coefp ? *F.coefp : F.coef) * Sv;
  const size_t g = F.glob[a];
  for (int i = 0; i < nc; ++i) t1[i + 3 * g] += (T)(F.valM[i + (size_t)F.dof * a] * sc);
#pragma unroll
  for (int c = 0; c < 3; ++c) t1s[1 + c + 4 * g] = W[c + (size_t)wdof * g] * t1[c + 3 * g];
}
// The face functional's coefficients q (FaceDev), one thread per support
// column, its pairs in their stored order, in FP64.
template <class T>
__global__ void face_q_kernel(int nsup, const int *__restrict__ qptr, const int *__restrict__ qa,
                              const int *__restrict__ qk, const int *__restrict__ glob,
                              const double *__restrict__ valM, int dof, MatView<T> G, double *__restrict__ q) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsup) return;
  const int nc = dof < 3 ? dof : 3;
  double v = 0.0;
  for (int p = qptr[s]; p < qptr[s + 1]; ++p) {
    const int a = qa[p], k = qk[p];
    const size_t g = glob[a];
    for (int i = 0; i < nc; ++i)
      v += valM[i + (size_t)dof * a] * (double)G.W[G.r0 + i + (size_t)G.wdof * g] *
           (double)G.base[(size_t)G.plane[4 * i] * G.nnz + k];
  }
  q[s] = v;
}
// DGKS decision: gate = ||w'||^2 < 0.5 ||w||^2
__global__ void dgks_kernel(const double *__restrict__ n0, const double *__restrict__ n1, int *__restrict__ gate,
                            const double *__restrict__ skip) {
  if (skip && *skip != 0.0) return;
  *gate = (*n1 < 0.5 * *n0) ? 1 : 0;
}
// AoS (vdof*vdof, nnz) -> planes, and Gt = -mG^T via the transpose map
template <class TS, class TD>
__global__ void aos_to_planes_kernel(size_t nnz, int m, const TS *__restrict__ aos, TD *__restrict__ V) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  for (int p = 0; p < m; ++p) V[(size_t)p * nnz + k] = (TD)aos[(size_t)m * k + p];
}
template <class TS, class TD>
__global__ void convert_kernel(size_t n, const TS *__restrict__ a, TD *__restrict__ b) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  b[k] = (TD)a[k];
}
template <class T>
__global__ void gt_kernel(size_t nnz, const T *__restrict__ V, const int *__restrict__ tr, T *__restrict__ Gt) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int l = tr[k];
  Gt[0 * nnz + l] = -V[3 * nnz + k];
  Gt[1 * nnz + l] = -V[7 * nnz + k];
  Gt[2 * nnz + l] = -V[11 * nnz + k];
}

// ----- CG (device-resident iterations with a convergence flag) -----
// scal layout: 0 err, 1 errO, 2 eps2, 3 den, 4 (unused), 5 done, 6 itr, 7 mItr, 8 suc
// Top-of-iteration test of CGRAD (memLS.f90:413-427): stop at mItr or when
// err < eps^2; otherwise errO = err and count the iteration.
__device__ inline void check_next(double *scal) {
  if (scal[5] != 0.0) return;
  if (scal[6] >= scal[7]) {
    scal[5] = 1.0;
    return;
  }
  if (scal[0] < scal[2]) {
    scal[8] = 1.0;
    scal[5] = 1.0;
    return;
  }
  scal[1] = scal[0];  // errO = err
  scal[6] += 1.0;
}
__global__ void cg_check_kernel(double *__restrict__ scal) { check_next(scal); }
// X += alpha P ; R -= alpha SP ; partial of R.R.  den = P . SP is summed
// from its nblk workgroup partials by every workgroup (reduce_kernel's
// order, so the same bits in each), which saves a launch per iteration.
template <class T>
__global__ void __launch_bounds__(256) cg_update_kernel(size_t L, const double *__restrict__ scal,
                                                        const T *__restrict__ dpart, int nblk,
                                                        T *__restrict__ X, T *__restrict__ R,
                                                        const T *__restrict__ P, const T *__restrict__ SP,
                                                        T *__restrict__ partial) {
  if (scal[5] != 0.0) return;
  __shared__ T sh[4];
  T d = T(0);
  for (int b = threadIdx.x; b < nblk; b += 256) d += dpart[b];
  const T den = block_sum_all(d, sh);
  const T alpha = (T)(scal[1] / (double)den);  // errO / den (CGRAD's alpha)
  const size_t base = (size_t)blockIdx.x * kChunk;
  T s = T(0);
#pragma unroll
  for (int t = 0; t < kChunk / 256; ++t) {
    const size_t e = base + threadIdx.x + 256 * t;
    if (e < L) {
      X[e] = X[e] + alpha * P[e];
      const T r = R[e] - alpha * SP[e];
      R[e] = r;
      s += r * r;
    }
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
// err = NORMS(R)**2 from the per-workgroup partials of R.R (fixed order);
// then beta = err/errO for the direction update, and the convergence test
// of the next iteration (cg_check_kernel's, with the same err).
template <class T>
__global__ void __launch_bounds__(256) cg_err_kernel(int nblk, const T *__restrict__ partial,
                                                     double *__restrict__ scal) {
  if (scal[5] != 0.0) return;
  __shared__ T sh[4];
  T s = T(0);
  for (int b = threadIdx.x; b < nblk; b += 256) s += partial[b];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    const double n = sqrt((double)s);
    scal[0] = n * n;
    scal[4] = scal[0] / scal[1];
    check_next(scal);
  }
}
// P = R + beta P, and optionally Ps = W(off + wdof i) P (the operator's
// scaled copy of the new direction)
template <class T>
__global__ void cg_p_kernel(size_t L, const double *__restrict__ scal, const T *__restrict__ R, T *__restrict__ P,
                            const T *__restrict__ W, int wdof, int off, T *__restrict__ Ps) {
  if (scal[5] != 0.0) return;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L) return;
  const T p = R[e] + (T)scal[4] * P[e];
  P[e] = p;
  if (Ps) Ps[e] = W[off + (size_t)wdof * e] * p;
}
}  // namespace

struct crimson_ls {
  SubLs GM, CG, RI;
  int LS_type = 0, Resm = 0, Resc = 0;
};

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The stored matrix planes and the -mG^T buffer in the solve's type.
inline const double *planes(const crimson_lhs *l, double) { return l->d_V.p; }
inline const float *planes(const crimson_lhs *l, float) { return l->d_V32.p; }
inline crimson::DBuf<double> &gt_buf(crimson_lhs *l, double) { return l->d_Gt; }
inline crimson::DBuf<float> &gt_buf(crimson_lhs *l, float) { return l->d_Gt32; }

// Per-solve context: device workspaces sized for this lhs and dof.
template <class T>
struct Ctx {
  crimson_lhs *lhs;
  hipStream_t s;
  int N, dof;
  crimson::DBuf<T> W, fix;
  crimson::DBuf<double> partial;  // per-workgroup partial sums (viewed as T for FP32 accumulation)
  crimson::DBuf<double> dsc, hcol;
  crimson::DBuf<double> fsum;  // one face dot per face (ADDBCMUL)
  crimson::DBuf<double> fpart;  // the Schur CG's face dot partials (kFaceDotBlocks per coupled face)
  crimson::DBuf<FaceDev> fdev;  // coupled faces with their PRE coefficients (Schur CG)
  std::vector<FaceDev> hfdev;
  crimson::DBuf<int> gate;
  std::vector<double> hbuf;
  bool coupled = false;
  int nblk_max = 0;
  // the CG's flag read-back: two pinned slots and their events
  double *hback = nullptr;
  hipEvent_t evb[2] = {nullptr, nullptr};

  Ctx() = default;
  Ctx(const Ctx &) = delete;
  Ctx &operator=(const Ctx &) = delete;
  ~Ctx() {
    if (hback || evb[0] || evb[1]) (void)hipStreamSynchronize(s);  // a read-back may still be in flight
    if (hback) (void)hipHostFree(hback);
    for (hipEvent_t e : evb)
      if (e) (void)hipEventDestroy(e);
  }
  int readback_slots() {
    if (!hback && hipHostMalloc((void **)&hback, 2 * 16 * sizeof(double)) != hipSuccess) {
      hback = nullptr;
      crimson::set_error("hipHostMalloc of the CG read-back slots failed");
      return CRIMSON_ENOMEM;
    }
    for (hipEvent_t &e : evb)
      if (!e) CR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return CRIMSON_OK;
  }

  int init(crimson_lhs *l, int d) {
    lhs = l;
    s = l->stream;
    N = l->nNo;
    dof = d;
    const size_t L = (size_t)dof * N;
    nblk_max = (int)((L + kChunk - 1) / kChunk);
    CR_TRY(W.alloc(L));
    CR_TRY(fix.alloc(L));
    CR_TRY(partial.alloc((size_t)nblk_max * 256));
    CR_TRY(dsc.alloc(64));
    CR_TRY(hcol.alloc(4 * 256));
    CR_TRY(fsum.alloc(std::max<size_t>(1, l->face.size())));
    CR_TRY(fpart.alloc(std::max<size_t>(1, l->face.size()) * kFaceDotBlocks));
    CR_TRY(gate.alloc(4));
    return CRIMSON_OK;
  }
};

template <class T>
MatView<T> view(const Ctx<T> &c, const T *base, int vdof, int r0, int c0, int nr, int nc, const T *W, const T *fix,
                int wdof) {
  MatView<T> A;
  A.base = base;
  A.nnz = (size_t)c.lhs->nnz;
  for (int k = 0; k < 16; ++k) A.plane[k] = 0;
  for (int r = 0; r < nr; ++r)
    for (int cc = 0; cc < nc; ++cc) A.plane[r * 4 + cc] = (r0 + r) * vdof + (c0 + cc);
  A.W = W;
  A.fix = fix;
  A.wdof = wdof;
  A.r0 = r0;
  A.c0 = c0;
  return A;
}

// y = ya * (A x) + za * z
template <class T>
int spmv(Ctx<T> &c, const MatView<T> &A, int nr, int nc, const T *x, T *y, double ya = 1.0, const T *z = nullptr,
         double za = 0.0, const double *skip = nullptr) {
  const unsigned TB = 256;
  const unsigned G = crimson::grid_for((size_t)c.N * kRowLanes, TB);
  const unsigned Gn = crimson::grid_for((size_t)c.N * kNarrowLanes, TB);
  const int *rp = c.lhs->d_rowptr.p, *cl = c.lhs->d_col.p;
  const short *c16 = c.lhs->d_col16.p;
#define CR_SPMV_IX(R_, C_, CL_)                                                                           \
  do {                                                                                                     \
    if (R_ * C_ >= 9)                                                                                      \
      spmv_kernel<R_, C_, T, kRowLanes><<<G, TB, 0, c.s>>>(c.N, rp, CL_, A, x, y, (T)ya, z, (T)za, skip);   \
    else                                                                                                   \
      spmv_kernel<R_, C_, T, kNarrowLanes><<<Gn, TB, 0, c.s>>>(c.N, rp, CL_, A, x, y, (T)ya, z, (T)za, skip); \
  } while (0)
#define CR_SPMV(R_, C_)                  \
  do {                                   \
    if (c16)                             \
      CR_SPMV_IX(R_, C_, c16);           \
    else                                 \
      CR_SPMV_IX(R_, C_, cl);            \
  } while (0)
  if (nr == 3 && nc == 3) CR_SPMV(3, 3);
  else if (nr == 4 && nc == 4) CR_SPMV(4, 4);
  else if (nr == 3 && nc == 1) CR_SPMV(3, 1);
  else if (nr == 1 && nc == 3) CR_SPMV(1, 3);
  else if (nr == 1 && nc == 1) CR_SPMV(1, 1);
  else if (nr == 2 && nc == 2) CR_SPMV(2, 2);
  else if (nr == 2 && nc == 1) CR_SPMV(2, 1);
  else if (nr == 1 && nc == 2) CR_SPMV(1, 2);
  else {
    crimson::set_error("spmv block %dx%d not instantiated", nr, nc);
    return CRIMSON_EUNSUPPORTED;
  }
#undef CR_SPMV
#undef CR_SPMV_IX
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

// dots of v with U_j (j < m) [+ v.v] -> dst on device (m [+1] values).
// Accumulation in T, or in FP64 when acc64 (the NS least-squares system and
// the norms that set the stopping tests of an FP32 solve).
template <class T>
int mdot(Ctx<T> &c, size_t L, const T *v, const T *U0, size_t ustride, int m, bool self, double *dst,
         const int *gate = nullptr, bool acc64 = false, const double *skip = nullptr) {
  const int nblk = (int)((L + kChunk - 1) / kChunk);
  const int pst = m + (self ? 1 : 0);
  if (pst > 256) {
    crimson::set_error("more than 255 simultaneous projections");
    return CRIMSON_EUNSUPPORTED;
  }
  if (acc64 || sizeof(T) == sizeof(double)) {
    mdot_kernel<T, double><<<nblk, 256, 0, c.s>>>(L, v, U0, ustride, m, self ? 1 : 0, c.partial.p, pst, gate,
                                                   skip);
    reduce_kernel<double><<<pst, 256, 0, c.s>>>(nblk, c.partial.p, pst, dst, gate, skip);
  } else {
    T *pt = reinterpret_cast<T *>(c.partial.p);
    mdot_kernel<T, T><<<nblk, 256, 0, c.s>>>(L, v, U0, ustride, m, self ? 1 : 0, pt, pst, gate, skip);
    reduce_kernel<T><<<pst, 256, 0, c.s>>>(nblk, pt, pst, dst, gate, skip);
  }
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

template <class T>
int axpby(Ctx<T> &c, size_t L, double a, const T *x, double b, T *y) {
  axpby_kernel<T><<<crimson::grid_for(L, 256), 256, 0, c.s>>>(L, (T)a, x, (T)b, y);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

template <class T>
int sync_read(Ctx<T> &c, const double *d, double *h, size_t n) {
  CR_HIP(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, c.s));
  CR_HIP(hipStreamSynchronize(c.s));
  return CRIMSON_OK;
}

template <class T>
double host_norm(Ctx<T> &c, size_t L, const T *v, int *rc) {
  *rc = mdot<T>(c, L, v, nullptr, 0, 0, true, c.dsc.p + 40, nullptr, true);
  if (*rc) return 0.0;
  double h = 0;
  *rc = sync_read(c, c.dsc.p + 40, &h, 1);
  return std::sqrt(h);
}

// ADDBCMUL (memLS.f90:206-268) over all coupled faces, X (xdof, N) -> Y.
// X may be Y: every face's dot product is taken (stream order) before any
// face's update is applied, which is what memLS's copy unCondU provides.
// skip (device, may be null): the device-driven solve's flag that makes
// these launches return at once; fcoef (device, may be null): every face's
// PRE coefficient computed on the device (indexed like lhs->face).
template <class T>
int addbcmul(Ctx<T> &c, bool pre, int xdof, const T *X, T *Y, const double *skip = nullptr,
             const double *fcoef = nullptr) {
  if (!c.coupled) return CRIMSON_OK;
  auto &faces = c.lhs->face;
  // all face dots first (PRE reads the unmodified vector, as unCondU does)
  for (size_t f = 0; f < faces.size(); ++f) {
    const CrFace &fa = faces[f];
    if (!fa.coupled) continue;
    face_dot_kernel<<<1, 1024, 0, c.s>>>(fa.nNo, xdof, fa.dof, fa.d_glob.p, fa.d_valM.p, X, c.fsum.p + f,
                                         skip);
  }
  for (size_t f = 0; f < faces.size(); ++f) {
    const CrFace &fa = faces[f];
    if (!fa.coupled) continue;
    if (fa.nNo == 0) continue;
    const double coef = pre ? -fa.res / (1.0 + fa.res * fa.nS) : fa.res;
    face_axpy_kernel<<<crimson::grid_for(fa.nNo, 256), 256, 0, c.s>>>(
        fa.nNo, xdof, fa.dof, fa.d_glob.p, fa.d_valM.p, c.fsum.p + f, coef, Y, (pre && fcoef) ? fcoef + f : nullptr,
        skip);
  }
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

// GMRES (memLS.f90:681-788) for the operator x -> A x + ADD(x), left
// preconditioned by PRE when faces are coupled; R, X are (nd, N) device.
template <class T>
struct Workspace {
  crimson::DBuf<T> U;  // (sD+1) vectors
  crimson::DBuf<T> tmp;
};

template <class T>
int gmres_dev(Ctx<T> &c, const MatView<T> &A, int nd, SubLs &ls, const T *R, T *X, Workspace<T> &ws, bool pre_ok,
              bool count_first = false) {
  const size_t L = (size_t)nd * c.N;
  const int sD = ls.sD;
  CR_TRY(ws.U.alloc(L * (sD + 1)));
  T *U = ws.U.p;
  std::vector<double> h((size_t)(sD + 1) * sD, 0.0), y(sD), cs(sD), sn(sD), err(sD + 1);
  auto H = [&](int i, int j) -> double & { return h[(size_t)i + (size_t)(sD + 1) * j]; };
  const double t0 = now_s();
  ls.suc = 0;
  CR_HIP(hipMemsetAsync(X, 0, L * sizeof(T), c.s));
  double eps = 0.0;
  int rc = CRIMSON_OK;
  std::vector<double> hc(2 * (sD + 2) + 4);
  for (int l = 1; l <= ls.mItr; ++l) {
    if (l == 1) {
      // X = 0 here, so R - A X is R (GMRESV still counts the product)
      if (count_first) ls.itr = ls.itr + 1;
      CR_HIP(hipMemcpyAsync(U, R, L * sizeof(T), hipMemcpyDeviceToDevice, c.s));
    } else {
      ls.itr = ls.itr + 1;
      CR_TRY(spmv(c, A, nd, nd, X, U));
      CR_TRY(addbcmul(c, false, nd, X, U));
      CR_TRY(axpby(c, L, 1.0, R, -1.0, U));
    }
    if (c.coupled && pre_ok) CR_TRY(addbcmul(c, true, nd, U, U));
    err[0] = host_norm(c, L, U, &rc);
    if (rc) return rc;
    if (l == 1) {
      eps = err[0];
      if (eps <= ls.absTol) {
        ls.callD = 0.0;
        ls.dB = 0.0;
        return CRIMSON_OK;
      }
      ls.iNorm = eps;
      ls.fNorm = eps;
      eps = std::max(ls.absTol, ls.relTol * eps);
    }
    ls.dB = ls.fNorm;
    CR_TRY(axpby(c, L, 1.0 / err[0], U, 0.0, U));
    int i;
    for (i = 1; i <= sD; ++i) {
      ls.itr = ls.itr + 1;
      T *ui = U + L * (size_t)i;
      const T *uprev = U + L * (size_t)(i - 1);
      CR_TRY(spmv(c, A, nd, nd, uprev, ui));
      CR_TRY(addbcmul(c, false, nd, uprev, ui));
      if (c.coupled && pre_ok) CR_TRY(addbcmul(c, true, nd, ui, ui));
      // CGS pass 1: h1 = U^T w, n0 = w.w ; w -= U h1 ; n1 = w.w
      double *h1 = c.hcol.p, *h2 = c.hcol.p + 256, *nrm = c.hcol.p + 512;
      const int nblk = (int)((L + kChunk - 1) / kChunk);
      CR_TRY(mdot(c, L, ui, U, L, i, true, h1));  // h1[i] = n0
      T *pt = reinterpret_cast<T *>(c.partial.p);
      msub_norm_kernel<T><<<nblk, 256, 0, c.s>>>(L, ui, U, L, i, h1, pt, 1, nullptr, nullptr);
      reduce_kernel<T><<<1, 256, 0, c.s>>>(nblk, pt, 1, nrm + 1, nullptr, nullptr);
      dgks_kernel<<<1, 1, 0, c.s>>>(h1 + i, nrm + 1, c.gate.p, nullptr);
      // DGKS pass 2 (gated)
      CR_TRY(mdot(c, L, ui, U, L, i, false, h2, c.gate.p));
      msub_norm_kernel<T><<<nblk, 256, 0, c.s>>>(L, ui, U, L, i, h2, pt, 1, c.gate.p, nullptr);
      reduce_kernel<T><<<1, 256, 0, c.s>>>(nblk, pt, 1, nrm + 2, c.gate.p, nullptr);
      CR_HIP(hipGetLastError());
      // read h1[0..i], h2[0..i-1], n1, n2, gate
      CR_HIP(hipMemcpyAsync(hc.data(), h1, (i + 1) * sizeof(double), hipMemcpyDeviceToHost, c.s));
      CR_HIP(hipMemcpyAsync(hc.data() + (sD + 2), h2, i * sizeof(double), hipMemcpyDeviceToHost, c.s));
      double nn[3];
      CR_HIP(hipMemcpyAsync(nn, nrm, 3 * sizeof(double), hipMemcpyDeviceToHost, c.s));
      int gate = 0;
      CR_HIP(hipMemcpyAsync(&gate, c.gate.p, sizeof(int), hipMemcpyDeviceToHost, c.s));
      CR_HIP(hipStreamSynchronize(c.s));
      for (int j = 1; j <= i; ++j) H(j - 1, i - 1) = hc[j - 1] + (gate ? hc[(sD + 2) + j - 1] : 0.0);
      H(i, i - 1) = std::sqrt(gate ? nn[2] : nn[1]);
      CR_TRY(axpby(c, L, 1.0 / H(i, i - 1), ui, 0.0, ui));
      for (int j = 1; j <= i - 1; ++j) {
        const double tmp = cs[j - 1] * H(j - 1, i - 1) + sn[j - 1] * H(j, i - 1);
        H(j, i - 1) = -sn[j - 1] * H(j - 1, i - 1) + cs[j - 1] * H(j, i - 1);
        H(j - 1, i - 1) = tmp;
      }
      const double tmp = std::sqrt(H(i - 1, i - 1) * H(i - 1, i - 1) + H(i, i - 1) * H(i, i - 1));
      cs[i - 1] = H(i - 1, i - 1) / tmp;
      sn[i - 1] = H(i, i - 1) / tmp;
      H(i - 1, i - 1) = tmp;
      H(i, i - 1) = 0.0;
      err[i] = -sn[i - 1] * err[i - 1];
      err[i - 1] = cs[i - 1] * err[i - 1];
      if (std::fabs(err[i]) < eps) {
        ls.suc = 1;
        break;
      }
    }
    if (i > sD) i = sD;
    for (int j = 0; j < i; ++j) y[j] = err[j];
    for (int j = i; j >= 1; --j) {
      for (int k = j + 1; k <= i; ++k) y[j - 1] = y[j - 1] - H(j - 1, k - 1) * y[k - 1];
      y[j - 1] = y[j - 1] / H(j - 1, j - 1);
    }
    CR_HIP(hipMemcpyAsync(c.hcol.p + 768, y.data(), i * sizeof(double), hipMemcpyHostToDevice, c.s));
    maxpy_kernel<T><<<crimson::grid_for(L, 256), 256, 0, c.s>>>(L, X, U, L, i, c.hcol.p + 768, 1.0);
    CR_HIP(hipGetLastError());
    CR_HIP(hipStreamSynchronize(c.s));  // y is a host buffer read by the copy above
    ls.fNorm = std::fabs(err[i]);
    if (ls.suc) break;
  }
  ls.callD = now_s() - t0 + ls.callD;
  ls.dB = 10.0 * std::log(ls.fNorm / ls.dB);
  return CRIMSON_OK;
}

// The operator's column-scaled copy of the CG direction, Ps(i) = W(off +
// wdof i) P(i), kept up to date by the direction update (Ps null: none).
template <class T>
struct Prescale {
  const T *W = nullptr;
  int wdof = 0, off = 0;
  T *Ps = nullptr;
};

// Generic CG driver on device vectors (memLS CGRAD / CGRADV / CGRADS).
// op(P, SP) computes SP for the iteration; R is overwritten with X.
template <class T, class Op>
int cg_dev(Ctx<T> &c, size_t L, SubLs &ls, T *R, Op op, T *X, T *P, T *SP, bool schur,
           Prescale<T> ps = Prescale<T>()) {
  const double t0 = now_s();
  double *scal = c.dsc.p;  // 0 err 1 errO 2 eps2 3 den 4 beta 5 done 6 itr 7 mItr 8 suc
  ls.suc = 0;
  int rc = CRIMSON_OK;
  const double errO0 = host_norm(c, L, R, &rc);
  if (rc) return rc;
  ls.iNorm = errO0;
  double eps = std::max(ls.absTol, ls.relTol * errO0);
  eps = eps * eps;
  double hs[9] = {errO0 * errO0, errO0 * errO0, eps, 0, 0, 0, 0, (double)ls.mItr, 0};
  CR_HIP(hipMemcpyAsync(scal, hs, sizeof(hs), hipMemcpyHostToDevice, c.s));
  CR_HIP(hipMemsetAsync(X, 0, L * sizeof(T), c.s));
  CR_HIP(hipMemcpyAsync(P, R, L * sizeof(T), hipMemcpyDeviceToDevice, c.s));
  if (ps.Ps)
    scale_p_kernel<T><<<crimson::grid_for(L, 256), 256, 0, c.s>>>((int)L, ps.W, ps.wdof, ps.off, P, ps.Ps, nullptr);
  const int nblk = (int)((L + kChunk - 1) / kChunk);
  const int batch = 8;
  double back[9];
  CR_TRY(c.readback_slots());
  // a batch of iterations, then the flags' copy into read-back slot `slot`
  auto enqueue_batch = [&](int slot) -> int {
    for (int b = 0; b < batch; ++b) {
      CR_TRY(op(P, SP, scal));
      // den = P . SP
      T *pt = reinterpret_cast<T *>(c.partial.p);
      mdot_kernel<T, T><<<nblk, 256, 0, c.s>>>(L, SP, P, 0, 1, 0, pt, 1, nullptr, nullptr);
      cg_update_kernel<T><<<nblk, 256, 0, c.s>>>(L, scal, pt, nblk, X, R, P, SP, pt + nblk);
      cg_err_kernel<T><<<1, 256, 0, c.s>>>(nblk, pt + nblk, scal);
      cg_p_kernel<T><<<crimson::grid_for(L, 256), 256, 0, c.s>>>(L, scal, R, P, ps.W, ps.wdof, ps.off, ps.Ps);
    }
    CR_HIP(hipGetLastError());
    CR_HIP(hipMemcpyAsync(c.hback + 16 * slot, scal, sizeof(back), hipMemcpyDeviceToHost, c.s));
    CR_HIP(hipEventRecord(c.evb[slot], c.s));
    return CRIMSON_OK;
  };
  // the first iteration's test; cg_err_kernel makes each later one
  cg_check_kernel<<<1, 1, 0, c.s>>>(scal);
  CR_TRY(enqueue_batch(0));
  for (int k = 0;; ++k) {
    // the next batch is queued before this one's flags are read, so that the
    // GPU does not idle while the host waits and launches; once the flag is
    // raised every kernel that changes the state returns at once, so the
    // extra batch leaves the state as this one did
    CR_TRY(enqueue_batch((k + 1) & 1));
    CR_HIP(hipEventSynchronize(c.evb[k & 1]));
    std::memcpy(back, c.hback + 16 * (k & 1), sizeof(back));
    if (back[5] != 0.0) break;
  }
  // the flag is raised at the top of the iteration after the last one
  CR_HIP(hipMemcpyAsync(R, X, L * sizeof(T), hipMemcpyDeviceToDevice, c.s));
  const double err = back[0], errO = back[1];
  ls.suc = back[8] != 0.0;
  ls.fNorm = std::sqrt(err);
  ls.dB = 5.0 * std::log(err / errO);
  const int done_itr = (int)back[6];
  if (schur)
    ls.itr = ls.itr + done_itr;  // CGRAD accumulates (memLS.f90:428)
  else
    ls.itr = done_itr;
  ls.callD = now_s() - t0 + (schur ? ls.callD : 0.0);
  return CRIMSON_OK;
}

// NSSOLVER (memLS.f90:1460-1759) on device; Ri (4,N) device in/out
// (preconditioned RHS in, preconditioned solution out).
// The support of a coupled face's dot functional (FaceDev): the columns of
// the face nodes' rows, each with its (face node, CSR entry) pairs, from the
// host copy of the pattern; built once per face (the pattern is fixed).
int face_support(crimson_lhs *l, CrFace &fa) {
  if (fa.qpat) return CRIMSON_OK;
  std::vector<std::array<int, 3>> t;  // (column j, face node a, entry k)
  for (int a = 0; a < fa.nNo; ++a) {
    const int g = fa.glob[a];
    for (int k = l->h_rowptr[g]; k < l->h_rowptr[g + 1]; ++k) t.push_back({l->h_col[k], a, k});
  }
  std::sort(t.begin(), t.end());
  std::vector<int> sup, ptr, qa, qk;
  for (size_t p = 0; p < t.size(); ++p) {
    if (p == 0 || t[p][0] != t[p - 1][0]) {
      sup.push_back(t[p][0]);
      ptr.push_back((int)p);
    }
    qa.push_back(t[p][1]);
    qk.push_back(t[p][2]);
  }
  ptr.push_back((int)t.size());
  fa.nsup = (int)sup.size();
  if (fa.nsup > 0) {
    CR_TRY(fa.d_sup.upload(sup.data(), sup.size(), l->stream));
    CR_TRY(fa.d_qptr.upload(ptr.data(), ptr.size(), l->stream));
    CR_TRY(fa.d_qa.upload(qa.data(), qa.size(), l->stream));
    CR_TRY(fa.d_qk.upload(qk.data(), qk.size(), l->stream));
    CR_TRY(fa.d_q.alloc(sup.size()));
    CR_HIP(hipStreamSynchronize(l->stream));
  }
  fa.qpat = true;
  return CRIMSON_OK;
}

template <class T>
int nssolver_dev(Ctx<T> &c, crimson_ls *ls, T *Ri) {
  crimson_lhs *l = c.lhs;
  const int N = c.N, nsd = 3;
  const size_t mn = (size_t)nsd * N;
  int iB = ls->RI.mItr;
  const int nB = 2 * iB;
  crimson::DBuf<T> Rm, Rc, Rmi, Rci, U, P, MU, MP, t1, t2, t3, t4;
  crimson::DBuf<double> dxB;
  CR_TRY(Rm.alloc(mn));
  CR_TRY(Rc.alloc(N));
  CR_TRY(Rmi.alloc(mn));
  CR_TRY(Rci.alloc(N));
  CR_TRY(U.alloc(mn * iB));
  CR_TRY(P.alloc((size_t)N * iB));
  CR_TRY(MU.alloc(mn * nB));
  CR_TRY(MP.alloc((size_t)N * nB));
  CR_TRY(t1.alloc(mn));
  CR_TRY(t2.alloc(N));
  CR_TRY(t3.alloc(N));
  CR_TRY(t4.alloc(N));
  CR_TRY(dxB.alloc(nB));
  Workspace<T> ws;
  CR_TRY(ws.tmp.alloc(std::max(mn, (size_t)4 * N)));  // GMRES scratch; the Schur CG's Z(4, N)
  const unsigned gN = crimson::grid_for(N, 256);
  extract_kernel<<<gN, 256, 0, c.s>>>(N, 4, 0, 3, Ri, Rmi.p);
  extract_kernel<<<gN, 256, 0, c.s>>>(N, 4, 3, 1, Ri, Rci.p);
  CR_HIP(hipMemcpyAsync(Rm.p, Rmi.p, mn * sizeof(T), hipMemcpyDeviceToDevice, c.s));
  CR_HIP(hipMemcpyAsync(Rc.p, Rci.p, N * sizeof(T), hipMemcpyDeviceToDevice, c.s));
  int rc = CRIMSON_OK;
  const double a1 = host_norm(c, mn, Rm.p, &rc);
  if (rc) return rc;
  const double a2 = host_norm(c, N, Rc.p, &rc);
  if (rc) return rc;
  double eps = std::sqrt(a1 * a1 + a2 * a2);
  ls->RI.iNorm = eps;
  ls->RI.fNorm = eps;
  ls->CG.callD = 0.0;
  ls->GM.callD = 0.0;
  ls->CG.itr = 0;
  ls->GM.itr = 0;
  ls->RI.suc = 0;
  const double t0 = now_s();
  eps = std::max(ls->RI.absTol, ls->RI.relTol * eps);
  // DEPART: views on the 4x4 planes (memLS.f90:1612-1679); Gt = -mG^T
  const T *V = planes(l, T(0));
  const MatView<T> mK = view<T>(c, V, 4, 0, 0, 3, 3, c.W.p, c.fix.p, 4);
  const MatView<T> mG = view<T>(c, V, 4, 0, 3, 3, 1, c.W.p, nullptr, 4);
  const MatView<T> mD = view<T>(c, V, 4, 3, 0, 1, 3, c.W.p, nullptr, 4);
  const MatView<T> mL = view<T>(c, V, 4, 3, 3, 1, 1, c.W.p, c.fix.p, 4);
  MatView<T> Gt = mD;
  if (!l->gt_is_d) {
    crimson::DBuf<T> &gtb = gt_buf(l, T(0));
    CR_TRY(gtb.alloc((size_t)3 * l->nnz));
    gt_kernel<T><<<crimson::grid_for(l->nnz, 256), 256, 0, c.s>>>(l->nnz, V, l->d_tr.p, gtb.p);
    CR_HIP(hipGetLastError());
    Gt = view<T>(c, gtb.p, 1, 0, 0, 1, 3, c.W.p, nullptr, 4);
    Gt.plane[0] = 0;
    Gt.plane[1] = 1;
    Gt.plane[2] = 2;
    Gt.r0 = 3;
    Gt.c0 = 0;
  }
  // BCPRE (memLS.f90:1683-1716): nS = ||valM||^2 over nsd components
  for (auto &fa : l->face) {
    if (!fa.coupled) continue;
    std::vector<double> vm((size_t)fa.dof * fa.nNo);
    CR_HIP(hipMemcpyAsync(vm.data(), fa.d_valM.p, vm.size() * sizeof(double), hipMemcpyDeviceToHost, c.s));
    CR_HIP(hipStreamSynchronize(c.s));
    fa.nS = 0.0;
    const int nc = std::min(fa.dof, nsd);
    for (int a = 0; a < fa.nNo; ++a)
      for (int i = 0; i < nc; ++i) fa.nS = fa.nS + vm[i + (size_t)fa.dof * a] * vm[i + (size_t)fa.dof * a];
  }
  // the coupled faces with their PRE coefficients and dot functionals
  c.hfdev.clear();
  for (auto &fa : l->face) {
    if (!fa.coupled) continue;
    CR_TRY(face_support(l, fa));
    if (fa.nsup > 0) {
      face_q_kernel<T><<<crimson::grid_for(fa.nsup, 256), 256, 0, c.s>>>(
          fa.nsup, fa.d_qptr.p, fa.d_qa.p, fa.d_qk.p, fa.d_glob.p, fa.d_valM.p, fa.dof, mG, fa.d_q.p);
      CR_HIP(hipGetLastError());
    }
    c.hfdev.push_back(FaceDev{fa.d_glob.p, fa.d_valM.p, -fa.res / (1.0 + fa.res * fa.nS), nullptr, fa.res, fa.nNo, fa.dof,
                              fa.d_sup.p, fa.d_q.p, fa.nsup});
  }
  const int nfc = (int)c.hfdev.size();
  if (nfc > 0) {
    CR_TRY(c.fdev.alloc(c.hfdev.size()));
    CR_HIP(hipMemcpyAsync(c.fdev.p, c.hfdev.data(), c.hfdev.size() * sizeof(FaceDev), hipMemcpyHostToDevice, c.s));
    CR_HIP(hipStreamSynchronize(c.s));
  }
  // Schur operator of CGRAD: SP = L P - Gt (G P + PRE), on P and its
  // scaled copy Ps = W_p o P in t2 (kept by the CG's direction update)
  auto schur_op = [&](const T *Pv, T *SPv, const double *scal) -> int {
    const double *done = scal + 5;
    const unsigned gr = crimson::grid_for((size_t)N * kNarrowLanes, 256);
    if (l->d_col16.p)
      g_product_kernel<T, short><<<gr + nfc * kFaceDotBlocks, 256, 0, c.s>>>(
          N, l->d_rowptr.p, l->d_col16.p, mG, t2.p, t1.p, ws.tmp.p, nfc, c.fdev.p, c.fpart.p, done);
    else
      g_product_kernel<T, int><<<gr + nfc * kFaceDotBlocks, 256, 0, c.s>>>(
          N, l->d_rowptr.p, l->d_col.p, mG, t2.p, t1.p, ws.tmp.p, nfc, c.fdev.p, c.fpart.p, done);
    for (int f = 0; f < nfc; ++f)
      if (c.hfdev[f].nNo > 0)
        face_pre_apply_kernel<T><<<crimson::grid_for(c.hfdev[f].nNo, 256), 256, 0, c.s>>>(
          c.hfdev[f], c.fpart.p + f * kFaceDotBlocks, t1.p, ws.tmp.p, c.W.p, 4, done);
    if (l->d_col16.p)
      schur_kernel<T, short><<<gr, 256, 0, c.s>>>(N, l->d_rowptr.p, l->d_col16.p, mL, Gt, Pv, t2.p, ws.tmp.p, SPv,
                                                  done);
    else
      schur_kernel<T, int><<<gr, 256, 0, c.s>>>(N, l->d_rowptr.p, l->d_col.p, mL, Gt, Pv, t2.p, ws.tmp.p, SPv, done);
    CR_HIP(hipGetLastError());
    return CRIMSON_OK;
  };
  Prescale<T> pre_p;
  pre_p.W = c.W.p;
  pre_p.wdof = 4;
  pre_p.off = mG.c0;
  pre_p.Ps = t2.p;
  std::vector<double> A((size_t)nB * nB, 0.0), Bv(nB, 0.0), xB(nB, 0.0);
  std::vector<double> dots(2 * nB + 4);
  int i, iBB = 0;
  for (i = 1; i <= ls->RI.mItr; ++i) {
    iB = 2 * i - 1;
    iBB = 2 * i;
    ls->RI.dB = ls->RI.fNorm;
    T *Ui = U.p + mn * (i - 1), *Pi = P.p + (size_t)N * (i - 1);
    T *MUb = MU.p + mn * (iB - 1), *MUbb = MU.p + mn * (iBB - 1);
    T *MPb = MP.p + (size_t)N * (iB - 1), *MPbb = MP.p + (size_t)N * (iBB - 1);
    CR_TRY(gmres_dev(c, mK, 3, ls->GM, Rm.p, Ui, ws, true));
    CR_TRY(spmv(c, mD, 1, 3, Ui, Pi, -1.0, Rc.p, 1.0));  // P = Rc - D U
    CR_TRY(cg_dev(c, (size_t)N, ls->CG, Pi, schur_op, t3.p, t4.p, MPb, true, pre_p));
    CR_TRY(spmv(c, mG, 3, 1, Pi, MUb));
    CR_TRY(axpby(c, mn, 1.0, Rm.p, 0.0, MUbb));
    CR_TRY(axpby(c, mn, -1.0, MUb, 1.0, MUbb));  // MU(iBB) = Rm - MU(iB)
    CR_TRY(gmres_dev(c, mK, 3, ls->GM, MUbb, Ui, ws, true));
    CR_TRY(spmv(c, mK, 3, 3, Ui, MUbb));
    CR_TRY(addbcmul(c, false, 3, Ui, MUbb));
    CR_TRY(spmv(c, mL, 1, 1, Pi, MPb));
    CR_TRY(spmv(c, mD, 1, 3, Ui, MPbb));
    // Gram matrix rows iB, iBB: dots with all previous columns and the RHS
    for (int k = iB; k <= iBB; ++k) {
      const T *MUk = MU.p + mn * (k - 1), *MPk = MP.p + (size_t)N * (k - 1);
      // MU_j . MU_k for j < k, then MU_k.MU_k ; plus MU_k . Rmi
      CR_TRY(mdot<T>(c, mn, MUk, MU.p, mn, k - 1, true, c.hcol.p, nullptr, true));
      CR_TRY(mdot<T>(c, mn, MUk, Rmi.p, 0, 1, false, c.hcol.p + 256, nullptr, true));
      CR_TRY(mdot<T>(c, N, MPk, MP.p, N, k - 1, true, c.hcol.p + 512, nullptr, true));
      CR_TRY(mdot<T>(c, N, MPk, Rci.p, 0, 1, false, c.hcol.p + 768, nullptr, true));
      std::vector<double> h1(k + 1), h2(k + 1), r1(1), r2(1);
      CR_HIP(hipMemcpyAsync(h1.data(), c.hcol.p, (k) * sizeof(double), hipMemcpyDeviceToHost, c.s));
      CR_HIP(hipMemcpyAsync(r1.data(), c.hcol.p + 256, sizeof(double), hipMemcpyDeviceToHost, c.s));
      CR_HIP(hipMemcpyAsync(h2.data(), c.hcol.p + 512, (k) * sizeof(double), hipMemcpyDeviceToHost, c.s));
      CR_HIP(hipMemcpyAsync(r2.data(), c.hcol.p + 768, sizeof(double), hipMemcpyDeviceToHost, c.s));
      CR_HIP(hipStreamSynchronize(c.s));
      for (int j = 1; j <= k - 1; ++j) {
        A[(j - 1) + (size_t)nB * (k - 1)] = h1[j - 1] + h2[j - 1];
        A[(k - 1) + (size_t)nB * (j - 1)] = A[(j - 1) + (size_t)nB * (k - 1)];
      }
      const double n1 = std::sqrt(h1[k - 1]), n2 = std::sqrt(h2[k - 1]);
      A[(k - 1) + (size_t)nB * (k - 1)] = n1 * n1 + n2 * n2;
      Bv[k - 1] = r1[0] + r2[0];
    }
    for (int k = 0; k < iBB; ++k) xB[k] = Bv[k];
    // GE (memLS.f90:607-677)
    {
      const int Nn = iBB;
      if (Nn == 2) {
        const double a11 = A[0], a21 = A[1], a12 = A[(size_t)nB], a22 = A[(size_t)nB + 1];
        const double pivot = a11 * a22 - a21 * a12;
        const double saveEl = (xB[0] * a22 - xB[1] * a12) / pivot;
        xB[1] = (xB[1] * a11 - xB[0] * a21) / pivot;
        xB[0] = saveEl;
      } else {
        std::vector<double> Cm((size_t)Nn * (Nn + 1));
        auto CC = [&](int r, int cc) -> double & { return Cm[r + (size_t)Nn * cc]; };
        for (int cc = 0; cc < Nn; ++cc)
          for (int r = 0; r < Nn; ++r) CC(r, cc) = A[r + (size_t)nB * cc];
        for (int r = 0; r < Nn; ++r) CC(r, Nn) = xB[r];
        for (int m = 0; m < Nn - 1; ++m) {
          int ipv = m;
          double pivot = std::fabs(CC(m, m));
          for (int r = m + 1; r < Nn; ++r)
            if (std::fabs(CC(r, m)) > pivot) {
              ipv = r;
              pivot = std::fabs(CC(r, m));
            }
          if (pivot < 2.0 * 2.220446049250313e-16) {
            crimson::set_error("memLS GE: singular matrix");
            return CRIMSON_ESINGULAR;
          }
          if (ipv != m)
            for (int cc = m; cc <= Nn; ++cc) std::swap(CC(m, cc), CC(ipv, cc));
          for (int r = m + 1; r < Nn; ++r) {
            const double sv = CC(r, m) / CC(m, m);
            CC(r, m) = 0.0;
            for (int cc = m + 1; cc <= Nn; ++cc) CC(r, cc) = CC(r, cc) - sv * CC(m, cc);
          }
        }
        for (int j = Nn - 1; j >= 0; --j) {
          for (int r = j + 1; r < Nn; ++r) CC(j, Nn) = CC(j, Nn) - CC(j, r) * CC(r, Nn);
          CC(j, Nn) = CC(j, Nn) / CC(j, j);
        }
        for (int r = 0; r < Nn; ++r) xB[r] = CC(r, Nn);
      }
    }
    double sxb = 0.0;
    for (int k = 0; k < iBB; ++k) sxb = sxb + xB[k] * Bv[k];
    ls->RI.fNorm = std::sqrt(ls->RI.iNorm * ls->RI.iNorm - sxb);
    if (ls->RI.fNorm < eps) {
      ls->RI.suc = 1;
      break;
    }
    CR_HIP(hipMemcpyAsync(dxB.p, xB.data(), iBB * sizeof(double), hipMemcpyHostToDevice, c.s));
    CR_TRY(axpby(c, mn, 1.0, Rmi.p, 0.0, Rm.p));
    CR_TRY(axpby(c, N, 1.0, Rci.p, 0.0, Rc.p));
    maxpy_kernel<<<crimson::grid_for(mn, 256), 256, 0, c.s>>>(mn, Rm.p, MU.p, mn, iBB, dxB.p, -1.0);
    maxpy_kernel<<<crimson::grid_for(N, 256), 256, 0, c.s>>>(N, Rc.p, MP.p, N, iBB, dxB.p, -1.0);
    CR_HIP(hipGetLastError());
    CR_HIP(hipStreamSynchronize(c.s));
  }
  if (i > ls->RI.mItr) {
    ls->RI.itr = ls->RI.mItr;
  } else {
    ls->RI.itr = i;
    CR_HIP(hipMemcpyAsync(dxB.p, xB.data(), iBB * sizeof(double), hipMemcpyHostToDevice, c.s));
    CR_TRY(axpby(c, N, 1.0, Rci.p, 0.0, Rc.p));
    maxpy_kernel<<<crimson::grid_for(N, 256), 256, 0, c.s>>>(N, Rc.p, MP.p, N, iBB, dxB.p, -1.0);
    CR_HIP(hipGetLastError());
  }
  {
    const double rcn = host_norm(c, N, Rc.p, &rc) / ls->RI.fNorm;
    if (rc) return rc;
    ls->Resc = (int)std::lround(100.0 * rcn * rcn);
    ls->Resm = 100 - ls->Resc;
  }
  // solution: Rmi = sum xB(2i) U_i, Rci = sum xB(2i-1) P_i
  {
    const int nit = ls->RI.itr;
    std::vector<double> cu(nit), cp(nit);
    for (int k = 1; k <= nit; ++k) {
      cu[k - 1] = xB[2 * k - 1];
      cp[k - 1] = xB[2 * k - 2];
    }
    CR_HIP(hipMemcpyAsync(c.hcol.p, cu.data(), nit * sizeof(double), hipMemcpyHostToDevice, c.s));
    CR_HIP(hipMemcpyAsync(c.hcol.p + 256, cp.data(), nit * sizeof(double), hipMemcpyHostToDevice, c.s));
    CR_HIP(hipMemsetAsync(Rmi.p, 0, mn * sizeof(T), c.s));
    CR_HIP(hipMemsetAsync(Rci.p, 0, N * sizeof(T), c.s));
    maxpy_kernel<<<crimson::grid_for(mn, 256), 256, 0, c.s>>>(mn, Rmi.p, U.p, mn, nit, c.hcol.p, 1.0);
    maxpy_kernel<<<crimson::grid_for(N, 256), 256, 0, c.s>>>(N, Rci.p, P.p, N, nit, c.hcol.p + 256, 1.0);
    insert_kernel<<<gN, 256, 0, c.s>>>(N, 4, 0, 3, Rmi.p, Ri);
    insert_kernel<<<gN, 256, 0, c.s>>>(N, 4, 3, 1, Rci.p, Ri);
    CR_HIP(hipGetLastError());
    CR_HIP(hipStreamSynchronize(c.s));
  }
  ls->RI.callD = now_s() - t0;
  ls->RI.dB = 10.0 * std::log(ls->RI.fNorm / ls->RI.dB);
  return CRIMSON_OK;
}

// One memLS_SOLVE in arithmetic type T.  d_Ri is the FP64 (dof,N) right-hand
// side in and solution out; for T = float the solve runs on an FP32 copy.
template <class T>
int solve_impl(crimson_lhs *l, crimson_ls *ls, int dof, double *d_Ri, const int *incL, const double *res) {
  Ctx<T> c;
  CR_TRY(c.init(l, dof));
  // face flags (memLS.f90:1965-1986)
  for (int f = 0; f < l->nFaces; ++f) {
    CrFace &fa = l->face[f];
    fa.inc = true;
    if (incL && incL[f] == 0) fa.inc = false;
    fa.coupled = false;
    if (!fa.inc || !fa.created) continue;
    if (fa.bGrp == CRIMSON_BC_NEU && res && res[f] != 0.0) {
      fa.res = res[f];
      fa.coupled = true;
      c.coupled = true;
    }
  }
  const int N = l->nNo;
  const size_t L = (size_t)dof * N;
  const unsigned gN = crimson::grid_for(N, 256);
  const T *V = planes(l, T(0));
  // PRECOND (memLS.f90:1763-1943) as the W/fix vectors
  precond_diag_kernel<T><<<gN, 256, 0, c.s>>>(N, dof, l->vdof, V, l->nnz, l->d_diag.p, c.W.p, c.fix.p);
  for (auto &fa : l->face)
    if (fa.created && fa.inc && fa.bGrp == CRIMSON_BC_DIR && fa.nNo > 0)
      precond_face_kernel<T><<<crimson::grid_for(fa.nNo, 256), 256, 0, c.s>>>(fa.nNo, dof, fa.dof, fa.d_glob.p,
                                                                              fa.d_val.p, c.W.p);
  precond_fix_kernel<T><<<gN, 256, 0, c.s>>>(N, dof, l->vdof, V, l->nnz, l->d_diag.p, c.W.p, c.fix.p);
  for (auto &fa : l->face)
    if (fa.coupled && fa.nNo > 0) {
      CR_TRY(fa.d_valM.alloc((size_t)fa.dof * fa.nNo));
      face_valm_kernel<T><<<crimson::grid_for(fa.nNo, 256), 256, 0, c.s>>>(fa.nNo, dof, fa.dof, fa.d_glob.p,
                                                                            fa.d_val.p, c.W.p, fa.d_valM.p);
    }
  // R = W R (into the working vector of type T)
  crimson::DBuf<T> Rw;
  T *R = nullptr;
  if (sizeof(T) == sizeof(double)) {
    R = reinterpret_cast<T *>(d_Ri);
  } else {
    CR_TRY(Rw.alloc(L));
    R = Rw.p;
  }
  mul_kernel<double, T, T><<<crimson::grid_for(L, 256), 256, 0, c.s>>>(L, d_Ri, c.W.p, R);
  CR_HIP(hipGetLastError());
  int rc = CRIMSON_OK;
  if (ls->LS_type == CRIMSON_LS_NS) {
    rc = nssolver_dev<T>(c, ls, R);
  } else if (ls->LS_type == CRIMSON_LS_GMRES) {
    // GMRESV (memLS.f90:792-883): no PRE application, restart from X
    Workspace<T> ws;
    CR_TRY(ws.tmp.alloc(L));
    crimson::DBuf<T> X;
    CR_TRY(X.alloc(L));
    const MatView<T> A = view<T>(c, V, dof, 0, 0, dof, dof, c.W.p, c.fix.p, dof);
    ls->RI.itr = 0;
    const double nr = host_norm(c, L, R, &rc);
    if (rc) return rc;
    if (nr <= ls->RI.absTol) {
      ls->RI.iNorm = nr;
      ls->RI.fNorm = nr;
      ls->RI.suc = 0;
      ls->RI.callD = 0.0;
      ls->RI.dB = 0.0;
      CR_HIP(hipMemsetAsync(R, 0, L * sizeof(T), c.s));
    } else {
      // GMRESV counts the residual product of every cycle (memLS.f90:827-829)
      rc = gmres_dev<T>(c, A, dof, ls->RI, R, X.p, ws, false, true);
      if (rc) return rc;
      CR_HIP(hipMemcpyAsync(R, X.p, L * sizeof(T), hipMemcpyDeviceToDevice, c.s));
    }
  } else if (ls->LS_type == CRIMSON_LS_CG) {
    const MatView<T> A = view<T>(c, V, dof, 0, 0, dof, dof, c.W.p, c.fix.p, dof);
    crimson::DBuf<T> X, P, SP;
    CR_TRY(X.alloc(L));
    CR_TRY(P.alloc(L));
    CR_TRY(SP.alloc(L));
    auto op = [&](const T *Pv, T *SPv, const double *) -> int { return spmv<T>(c, A, dof, dof, Pv, SPv); };
    rc = cg_dev<T>(c, L, ls->RI, R, op, X.p, P.p, SP.p, false);
  } else {
    crimson::set_error("LS_type %d not defined", ls->LS_type);
    return CRIMSON_EARG;
  }
  if (rc) return rc;
  mul_kernel<T, T, double><<<crimson::grid_for(L, 256), 256, 0, c.s>>>(L, R, c.W.p, d_Ri);  // R = R W
  CR_HIP(hipGetLastError());
  CR_HIP(hipStreamSynchronize(c.s));
  return CRIMSON_OK;
}

int solve_dev(crimson_lhs *l, crimson_ls *ls, int dof, double *d_Ri, const int *incL, const double *res) {
  if (l->vdof != dof || !(l->prec == 32 ? (const void *)l->d_V32.p : (const void *)l->d_V.p)) {
    crimson::set_error("lhs holds no matrix of block size %d (stored: %d)", dof, l->vdof);
    return CRIMSON_ESTATE;
  }
  if (ls->LS_type == CRIMSON_LS_NS && dof != 4) {
    crimson::set_error("LS_TYPE_NS needs dof = 4 (got %d)", dof);
    return CRIMSON_EUNSUPPORTED;
  }
  if (l->prec == 32) return solve_impl<float>(l, ls, dof, d_Ri, incL, res);
  return solve_impl<double>(l, ls, dof, d_Ri, incL, res);
}


// ---------------------------------------------------------------------------
// Device-driven NSSOLVER: the same kernels and the same arithmetic as
// nssolver_dev, with every host decision of the solve (GMRES's Givens
// rotations, back substitution and convergence test, CGRAD's iteration
// count, NSSOLVER's least-squares Gram system, its pivoted GE and stopping
// test, BCPRE's nS and the PRE coefficients) moved into one-thread kernels,
// so that the whole solve is one fixed launch sequence that a HIP graph can
// hold.  Every possible iteration is enqueued (GMRES: mItr cycles of sD
// Arnoldi steps; CGRAD: mItr iterations; NSSOLVER: mItr outer iterations);
// the launches past convergence see their flag and return at once.  The
// scalar kernels keep the host's operation order with FMA contraction off,
// so the solve gives the host-driven solve's bits.
//
// gst (double, device): the state of the solve and of its last GMRES and
// CG calls, then the GMRES Hessenberg system, the Gram system and the
// solution coefficients.
enum GstIdx : int {
  G_RI_SKIP = 0,  // the outer loop has stopped (converged or failed)
  G_RI_ITR,
  G_RI_SUC,
  G_RI_INORM,
  G_RI_FNORM,
  G_RI_EPS,
  G_RI_DBPRE,
  G_ERR,  // 1: singular GE
  G_GM_CYC = 8,  // this GMRES call does nothing more (done, or not running)
  G_GM_ARN,      // no further Arnoldi step in this cycle
  G_GM_EPS,
  G_GM_I,
  G_GM_ITR,
  G_GM_SUC,
  G_GM_INORM,
  G_GM_FNORM,
  G_GM_DBPRE,
  G_GM_SCALE,
  G_GM_APPLY,
  G_GM_RAN,
  G_GM_EARLY,
  G_GM_LASTF,  // fNorm and dB of the last GMRES call that ran to the end
  G_GM_LASTD,
  G_CG_ITR = 24,
  G_CG_SUC,
  G_CG_INORM,
  G_CG_ERR,
  G_CG_ERRO,
  G_RC_N2 = 30,
  G_NS_CNT,
  G_ARR = 64
};

template <class T>
__global__ void gfill_kernel(size_t L, T *__restrict__ x, const double *__restrict__ skip) {
  if (skip && *skip != 0.0) return;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < L) x[e] = T(0);
}
// dst = src when *skip is 0 (skip may be null) and *runif is not 0 (runif may be null)
template <class T>
__global__ void gcopy_kernel(size_t L, const T *__restrict__ src, T *__restrict__ dst,
                             const double *__restrict__ skip, const double *__restrict__ runif) {
  if ((skip && *skip != 0.0) || (runif && *runif == 0.0)) return;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < L) dst[e] = src[e];
}
// x = (T)(*s) x, as axpby(s, x, 0, x)
template <class T>
__global__ void gscale_kernel(size_t L, T *__restrict__ x, const double *__restrict__ s,
                              const double *__restrict__ skip) {
  if (skip && *skip != 0.0) return;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < L) x[e] = (T)(*s) * x[e];
}
template <class T>
__global__ void gaxpby_kernel(size_t L, T a, const T *__restrict__ x, T b, T *__restrict__ y,
                              const double *__restrict__ skip) {
  if (skip && *skip != 0.0) return;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L) return;
  y[e] = a * x[e] + (b == T(0) ? T(0) : b * y[e]);
}
// maxpy_kernel with the count m_host, or *m_dev when given
template <class T>
__global__ void gmaxpy_kernel(size_t L, T *__restrict__ y, const T *__restrict__ U0, size_t ustride, int m_host,
                              const double *__restrict__ m_dev, const double *__restrict__ coef, double sign,
                              const double *__restrict__ skip, const double *__restrict__ runif) {
  if ((skip && *skip != 0.0) || (runif && *runif == 0.0)) return;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L) return;
  const int m = m_dev ? (int)*m_dev : m_host;
  T w = y[e];
  for (int j = 0; j < m; ++j) w += (T)(sign * coef[j]) * U0[ustride * j + e];
  y[e] = w;
}

__device__ inline double dmax(double a, double b) { return a < b ? b : a; }  // std::max's choice

__global__ void ns_begin_kernel(double *__restrict__ g, const double *__restrict__ n2a,
                                const double *__restrict__ n2b, double absTol, double relTol) {
#pragma clang fp contract(off)
  const double a1 = sqrt(*n2a), a2 = sqrt(*n2b);
  const double eps = sqrt(a1 * a1 + a2 * a2);
  g[G_RI_INORM] = eps;
  g[G_RI_FNORM] = eps;
  g[G_RI_EPS] = dmax(absTol, relTol * eps);
  g[G_RI_SKIP] = 0.0;
  g[G_RI_SUC] = 0.0;
  g[G_RI_ITR] = 0.0;
  g[G_ERR] = 0.0;
  g[G_GM_ITR] = 0.0;
  g[G_CG_ITR] = 0.0;
  g[G_GM_LASTF] = g[G_GM_LASTD] = 0.0;
}
// BCPRE's nS of a coupled face (memLS.f90:1683-1716, summed in the host's
// order) and its PRE coefficient -res / (1 + res nS)
__global__ void face_pre_coef_kernel(int nNo, int fdof, int nsd, const double *__restrict__ valM, double res,
                                     double *__restrict__ coef) {
#pragma clang fp contract(off)
  double nS = 0.0;
  const int nc = fdof < nsd ? fdof : nsd;
  for (int a = 0; a < nNo; ++a)
    for (int i = 0; i < nc; ++i) nS = nS + valM[i + (size_t)fdof * a] * valM[i + (size_t)fdof * a];
  *coef = -res / (1.0 + res * nS);
}
__global__ void gm_begin_kernel(double *__restrict__ g) {
  if (g[G_RI_SKIP] != 0.0) {
    g[G_GM_CYC] = 1.0;
    g[G_GM_ARN] = 1.0;
    g[G_GM_RAN] = 0.0;
    return;
  }
  g[G_GM_CYC] = 0.0;
  g[G_GM_ARN] = 1.0;
  g[G_GM_SUC] = 0.0;
  g[G_GM_RAN] = 0.0;
  g[G_GM_EARLY] = 0.0;
}
__global__ void gm_count_kernel(double *__restrict__ g) {
  if (g[G_GM_CYC] != 0.0) return;
  g[G_GM_ITR] = g[G_GM_ITR] + 1.0;
}
__global__ void gm_cycle_kernel(double *__restrict__ g, int l, double absTol, double relTol,
                                const double *__restrict__ n2, double *__restrict__ err) {
#pragma clang fp contract(off)
  if (g[G_GM_CYC] != 0.0) return;
  const double e0 = sqrt(*n2);
  err[0] = e0;
  if (l == 1) {
    const double eps = e0;
    if (eps <= absTol) {
      g[G_GM_EARLY] = 1.0;
      g[G_GM_CYC] = 1.0;
      g[G_GM_ARN] = 1.0;
      return;
    }
    g[G_GM_INORM] = eps;
    g[G_GM_FNORM] = eps;
    g[G_GM_EPS] = dmax(absTol, relTol * eps);
  }
  g[G_GM_DBPRE] = g[G_GM_FNORM];
  g[G_GM_SCALE] = 1.0 / e0;
  g[G_GM_ARN] = 0.0;
  g[G_GM_RAN] = 1.0;
  g[G_GM_I] = 0.0;
}
// Arnoldi step i's Hessenberg column, rotations and test (gmres_dev's host
// part); H (sD+1, sD) column-major
__global__ void gm_givens_kernel(double *__restrict__ g, int i, int sD, const double *__restrict__ h1,
                                 const double *__restrict__ h2, const double *__restrict__ nrm,
                                 const int *__restrict__ gate, double *__restrict__ H, double *__restrict__ cs,
                                 double *__restrict__ sn, double *__restrict__ err) {
#pragma clang fp contract(off)
  if (g[G_GM_ARN] != 0.0) return;
  g[G_GM_ITR] = g[G_GM_ITR] + 1.0;
  const int ld = sD + 1;
  const bool gt = *gate != 0;
  for (int j = 1; j <= i; ++j) H[(j - 1) + (size_t)ld * (i - 1)] = h1[j - 1] + (gt ? h2[j - 1] : 0.0);
  H[i + (size_t)ld * (i - 1)] = sqrt(gt ? nrm[2] : nrm[1]);
  g[G_GM_SCALE] = 1.0 / H[i + (size_t)ld * (i - 1)];
  for (int j = 1; j <= i - 1; ++j) {
    const double a = H[(j - 1) + (size_t)ld * (i - 1)], b = H[j + (size_t)ld * (i - 1)];
    const double tmp = cs[j - 1] * a + sn[j - 1] * b;
    H[j + (size_t)ld * (i - 1)] = -sn[j - 1] * a + cs[j - 1] * b;
    H[(j - 1) + (size_t)ld * (i - 1)] = tmp;
  }
  const double a = H[(i - 1) + (size_t)ld * (i - 1)], b = H[i + (size_t)ld * (i - 1)];
  const double tmp = sqrt(a * a + b * b);
  cs[i - 1] = a / tmp;
  sn[i - 1] = b / tmp;
  H[(i - 1) + (size_t)ld * (i - 1)] = tmp;
  H[i + (size_t)ld * (i - 1)] = 0.0;
  err[i] = -sn[i - 1] * err[i - 1];
  err[i - 1] = cs[i - 1] * err[i - 1];
  g[G_GM_I] = (double)i;
  if (fabs(err[i]) < g[G_GM_EPS]) {
    g[G_GM_SUC] = 1.0;
    g[G_GM_ARN] = 1.0;
  }
}
// the cycle's back substitution (gmres_dev's host part)
__global__ void gm_solve_kernel(double *__restrict__ g, int sD, const double *__restrict__ H,
                                const double *__restrict__ err, double *__restrict__ y) {
#pragma clang fp contract(off)
  if (g[G_GM_CYC] != 0.0) return;
  const int ld = sD + 1;
  const int i = (int)g[G_GM_I];
  for (int j = 0; j < i; ++j) y[j] = err[j];
  for (int j = i; j >= 1; --j) {
    for (int k = j + 1; k <= i; ++k) y[j - 1] = y[j - 1] - H[(j - 1) + (size_t)ld * (k - 1)] * y[k - 1];
    y[j - 1] = y[j - 1] / H[(j - 1) + (size_t)ld * (j - 1)];
  }
  g[G_GM_APPLY] = (double)i;
  g[G_GM_FNORM] = fabs(err[i]);
}
__global__ void gm_endcycle_kernel(double *__restrict__ g) {
  if (g[G_GM_CYC] != 0.0) return;
  if (g[G_GM_SUC] != 0.0) g[G_GM_CYC] = 1.0;
  g[G_GM_ARN] = 1.0;
}
__global__ void gm_finish_kernel(double *__restrict__ g) {
  if (g[G_RI_SKIP] != 0.0 || g[G_GM_RAN] == 0.0) return;
  g[G_GM_LASTF] = g[G_GM_FNORM];
  g[G_GM_LASTD] = g[G_GM_DBPRE];
}
__global__ void cg_begin_kernel(double *__restrict__ scal, double *__restrict__ g, const double *__restrict__ n2,
                                double absTol, double relTol, int mItr) {
#pragma clang fp contract(off)
  if (g[G_RI_SKIP] != 0.0) {
    scal[5] = 1.0;
    return;
  }
  const double e = sqrt(*n2);
  g[G_CG_INORM] = e;
  double eps = dmax(absTol, relTol * e);
  eps = eps * eps;
  scal[0] = e * e;
  scal[1] = e * e;
  scal[2] = eps;
  scal[3] = scal[4] = scal[5] = scal[6] = 0.0;
  scal[7] = (double)mItr;
  scal[8] = 0.0;
}
__global__ void cg_end_kernel(const double *__restrict__ scal, double *__restrict__ g) {
  if (g[G_RI_SKIP] != 0.0) return;
  g[G_CG_ITR] = g[G_CG_ITR] + (double)(int)scal[6];
  g[G_CG_SUC] = scal[8];
  g[G_CG_ERR] = scal[0];
  g[G_CG_ERRO] = scal[1];
}
// NSSOLVER iteration i's Gram rows iB = 2i-1, iBB = 2i from their dots
// (dk: for each of the two, h1 at 0, r1 at 256, h2 at 512, r2 at 768),
// then GE (memLS.f90:607-677) and the stopping test; A (nB, nB), C the GE
// scratch (Nn, Nn+1)
__global__ void ns_gram_kernel(double *__restrict__ g, int i, int nB, const double *__restrict__ dk,
                               double *__restrict__ A, double *__restrict__ Bv, double *__restrict__ xB,
                               double *__restrict__ C) {
#pragma clang fp contract(off)
  if (g[G_RI_SKIP] != 0.0) return;
  const int iB = 2 * i - 1, iBB = 2 * i;
  g[G_RI_DBPRE] = g[G_RI_FNORM];
  for (int k = iB; k <= iBB; ++k) {
    const double *d = dk + (size_t)1024 * (k - iB);
    const double *h1 = d, *r1 = d + 256, *h2 = d + 512, *r2 = d + 768;
    for (int j = 1; j <= k - 1; ++j) {
      A[(j - 1) + (size_t)nB * (k - 1)] = h1[j - 1] + h2[j - 1];
      A[(k - 1) + (size_t)nB * (j - 1)] = A[(j - 1) + (size_t)nB * (k - 1)];
    }
    const double n1 = sqrt(h1[k - 1]), n2 = sqrt(h2[k - 1]);
    A[(k - 1) + (size_t)nB * (k - 1)] = n1 * n1 + n2 * n2;
    Bv[k - 1] = r1[0] + r2[0];
  }
  for (int k = 0; k < iBB; ++k) xB[k] = Bv[k];
  const int Nn = iBB;
  if (Nn == 2) {
    const double a11 = A[0], a21 = A[1], a12 = A[nB], a22 = A[nB + 1];
    const double pivot = a11 * a22 - a21 * a12;
    const double saveEl = (xB[0] * a22 - xB[1] * a12) / pivot;
    xB[1] = (xB[1] * a11 - xB[0] * a21) / pivot;
    xB[0] = saveEl;
  } else {
#define CC(r, cc) C[(r) + (size_t)Nn * (cc)]
    for (int cc = 0; cc < Nn; ++cc)
      for (int r = 0; r < Nn; ++r) CC(r, cc) = A[r + (size_t)nB * cc];
    for (int r = 0; r < Nn; ++r) CC(r, Nn) = xB[r];
    for (int m = 0; m < Nn - 1; ++m) {
      int ipv = m;
      double pivot = fabs(CC(m, m));
      for (int r = m + 1; r < Nn; ++r)
        if (fabs(CC(r, m)) > pivot) {
          ipv = r;
          pivot = fabs(CC(r, m));
        }
      if (pivot < 2.0 * 2.220446049250313e-16) {
        g[G_ERR] = 1.0;
        g[G_RI_SKIP] = 1.0;
        return;
      }
      if (ipv != m)
        for (int cc = m; cc <= Nn; ++cc) {
          const double t = CC(m, cc);
          CC(m, cc) = CC(ipv, cc);
          CC(ipv, cc) = t;
        }
      for (int r = m + 1; r < Nn; ++r) {
        const double sv = CC(r, m) / CC(m, m);
        CC(r, m) = 0.0;
        for (int cc = m + 1; cc <= Nn; ++cc) CC(r, cc) = CC(r, cc) - sv * CC(m, cc);
      }
    }
    for (int j = Nn - 1; j >= 0; --j) {
      for (int r = j + 1; r < Nn; ++r) CC(j, Nn) = CC(j, Nn) - CC(j, r) * CC(r, Nn);
      CC(j, Nn) = CC(j, Nn) / CC(j, j);
    }
    for (int r = 0; r < Nn; ++r) xB[r] = CC(r, Nn);
#undef CC
  }
  double sxb = 0.0;
  for (int k = 0; k < iBB; ++k) sxb = sxb + xB[k] * Bv[k];
  const double iN = g[G_RI_INORM];
  g[G_RI_FNORM] = sqrt(iN * iN - sxb);
  if (g[G_RI_FNORM] < g[G_RI_EPS]) {
    g[G_RI_SUC] = 1.0;
    g[G_RI_ITR] = (double)i;
    g[G_RI_SKIP] = 1.0;
  }
}
// after the outer loop: the iteration count, and the solution's
// coefficients cu(k) = xB(2k), cp(k) = xB(2k-1) (1-based)
__global__ void ns_final_kernel(double *__restrict__ g, int mItr, const double *__restrict__ xB,
                                double *__restrict__ cu, double *__restrict__ cp) {
  if (g[G_ERR] != 0.0) return;
  if (g[G_RI_SUC] == 0.0) g[G_RI_ITR] = (double)mItr;
  const int nit = (int)g[G_RI_ITR];
  g[G_NS_CNT] = 2.0 * nit;
  for (int k = 1; k <= nit; ++k) {
    cu[k - 1] = xB[2 * k - 1];
    cp[k - 1] = xB[2 * k - 2];
  }
}

// Device workspaces of one device-driven solve (allocated before capture).
template <class T>
struct NsGraphWork {
  Ctx<T> c;
  crimson::DBuf<T> Rm, Rc, Rmi, Rci, U, P, MU, MP, t1, t2, t3, t4, Rw, gmU, gmTmp;
  crimson::DBuf<double> gst, fcoef, dots;
  int sD = 0, nB = 0, mRI = 0, mGM = 0, mCG = 0;
  double riAbs = 0, riRel = 0, gmAbs = 0, gmRel = 0, cgAbs = 0, cgRel = 0;
  bool disjoint = false;  // the coupled faces share no node: one launch for all of them
  int maxfn = 0;          // the largest coupled face
  crimson::DBuf<double> msum;  // the faces' dots of the batched ADDBCMUL
  struct SegCapture *cap = nullptr;  // the capture in progress, cut between the solve's phases
  int cut();
  size_t oH = 0, ocs = 0, osn = 0, oerr = 0, oy = 0, oA = 0, oBv = 0, oxB = 0, oC = 0, ocu = 0, ocp = 0;
  double *G(size_t off) { return gst.p + off; }
};

// ADDBCMUL of the device-driven solve: the batched face kernels when the
// coupled faces are disjoint, addbcmul's per-face launches otherwise.
template <class T>
int addbcmul_graph(NsGraphWork<T> &w, bool pre, int xdof, const T *X, T *Y, const double *skip) {
  Ctx<T> &c = w.c;
  if (!c.coupled) return CRIMSON_OK;
  const int nfc = (int)c.hfdev.size();
  if (!w.disjoint || nfc < 2) return addbcmul(c, pre, xdof, X, Y, skip, pre ? w.fcoef.p : nullptr);
  face_dot_multi_kernel<T><<<nfc, 1024, 0, c.s>>>(c.fdev.p, xdof, X, w.msum.p, skip);
  face_axpy_multi_kernel<T><<<dim3(crimson::grid_for(std::max(1, w.maxfn), 256), nfc), 256, 0, c.s>>>(
      c.fdev.p, xdof, w.msum.p, pre ? 1 : 0, Y, skip);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

// GMRES (gmres_dev) device-driven, on the operator x -> A x + ADD(x) with
// PRE (the NS solver's momentum solves: pre_ok, GMRESV's first-product
// count not kept)
template <class T>
int gmres_graph(NsGraphWork<T> &w, const MatView<T> &A, int nd, const T *R, T *X) {
  Ctx<T> &c = w.c;
  const size_t L = (size_t)nd * c.N;
  const int sD = w.sD;
  double *g = w.gst.p;
  const double *cyc = g + G_GM_CYC, *arn = g + G_GM_ARN, *ri = g + G_RI_SKIP;
  T *U = w.gmU.p;
  const unsigned GL = crimson::grid_for(L, 256);
  gm_begin_kernel<<<1, 1, 0, c.s>>>(g);
  gfill_kernel<T><<<GL, 256, 0, c.s>>>(L, X, ri);
  const int nblk = (int)((L + kChunk - 1) / kChunk);
  double *h1 = c.hcol.p, *h2 = c.hcol.p + 256, *nrm = c.hcol.p + 512;
  T *pt = reinterpret_cast<T *>(c.partial.p);
  for (int l = 1; l <= w.mGM; ++l) {
    if (l == 1) {
      gcopy_kernel<T><<<GL, 256, 0, c.s>>>(L, R, U, cyc, nullptr);
    } else {
      gm_count_kernel<<<1, 1, 0, c.s>>>(g);
      CR_TRY(spmv(c, A, nd, nd, X, U, 1.0, static_cast<const T *>(nullptr), 0.0, cyc));
      CR_TRY(addbcmul_graph(w, false, nd, X, U, cyc));
      gaxpby_kernel<T><<<GL, 256, 0, c.s>>>(L, T(1), R, T(-1), U, cyc);
    }
    CR_TRY(addbcmul_graph(w, true, nd, U, U, cyc));
    CR_TRY(mdot<T>(c, L, U, nullptr, 0, 0, true, c.dsc.p + 40, nullptr, true, cyc));
    gm_cycle_kernel<<<1, 1, 0, c.s>>>(g, l, w.gmAbs, w.gmRel, c.dsc.p + 40, w.G(w.oerr));
    gscale_kernel<T><<<GL, 256, 0, c.s>>>(L, U, g + G_GM_SCALE, cyc);
    for (int i = 1; i <= sD; ++i) {
      T *ui = U + L * (size_t)i;
      const T *uprev = U + L * (size_t)(i - 1);
      CR_TRY(spmv(c, A, nd, nd, uprev, ui, 1.0, static_cast<const T *>(nullptr), 0.0, arn));
      CR_TRY(addbcmul_graph(w, false, nd, uprev, ui, arn));
      CR_TRY(addbcmul_graph(w, true, nd, ui, ui, arn));
      CR_TRY(mdot(c, L, ui, U, L, i, true, h1, nullptr, false, arn));
      msub_norm_kernel<T><<<nblk, 256, 0, c.s>>>(L, ui, U, L, i, h1, pt, 1, nullptr, arn);
      reduce_kernel<T><<<1, 256, 0, c.s>>>(nblk, pt, 1, nrm + 1, nullptr, arn);
      dgks_kernel<<<1, 1, 0, c.s>>>(h1 + i, nrm + 1, c.gate.p, arn);
      CR_TRY(mdot(c, L, ui, U, L, i, false, h2, c.gate.p, false, arn));
      msub_norm_kernel<T><<<nblk, 256, 0, c.s>>>(L, ui, U, L, i, h2, pt, 1, c.gate.p, arn);
      reduce_kernel<T><<<1, 256, 0, c.s>>>(nblk, pt, 1, nrm + 2, c.gate.p, arn);
      gm_givens_kernel<<<1, 1, 0, c.s>>>(g, i, sD, h1, h2, nrm, c.gate.p, w.G(w.oH), w.G(w.ocs), w.G(w.osn),
                                         w.G(w.oerr));
      gscale_kernel<T><<<GL, 256, 0, c.s>>>(L, ui, g + G_GM_SCALE, arn);
    }
    gm_solve_kernel<<<1, 1, 0, c.s>>>(g, sD, w.G(w.oH), w.G(w.oerr), w.G(w.oy));
    gmaxpy_kernel<T><<<GL, 256, 0, c.s>>>(L, X, U, L, 0, g + G_GM_APPLY, w.G(w.oy), 1.0, cyc, nullptr);
    gm_endcycle_kernel<<<1, 1, 0, c.s>>>(g);
  }
  gm_finish_kernel<<<1, 1, 0, c.s>>>(g);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

// CGRAD on the Schur operator (cg_dev with schur = true) device-driven:
// mItr iteration bodies, each launch returning at once after the flag.
template <class T, class Op>
int cg_graph(NsGraphWork<T> &w, size_t L, T *R, Op op, T *X, T *P, T *SP, Prescale<T> ps) {
  Ctx<T> &c = w.c;
  double *g = w.gst.p, *scal = c.dsc.p;
  const unsigned GL = crimson::grid_for(L, 256);
  CR_TRY(mdot<T>(c, L, R, nullptr, 0, 0, true, c.dsc.p + 41, nullptr, true, g + G_RI_SKIP));
  cg_begin_kernel<<<1, 1, 0, c.s>>>(scal, g, c.dsc.p + 41, w.cgAbs, w.cgRel, w.mCG);
  gfill_kernel<T><<<GL, 256, 0, c.s>>>(L, X, scal + 5);
  gcopy_kernel<T><<<GL, 256, 0, c.s>>>(L, R, P, scal + 5, nullptr);
  if (ps.Ps) scale_p_kernel<T><<<GL, 256, 0, c.s>>>((int)L, ps.W, ps.wdof, ps.off, P, ps.Ps, scal + 5);
  const int nblk = (int)((L + kChunk - 1) / kChunk);
  cg_check_kernel<<<1, 1, 0, c.s>>>(scal);
  for (int it = 0; it < w.mCG; ++it) {
    CR_TRY(op(P, SP, scal));
    T *pt = reinterpret_cast<T *>(c.partial.p);
    mdot_kernel<T, T><<<nblk, 256, 0, c.s>>>(L, SP, P, 0, 1, 0, pt, 1, nullptr, scal + 5);
    cg_update_kernel<T><<<nblk, 256, 0, c.s>>>(L, scal, pt, nblk, X, R, P, SP, pt + nblk);
    cg_err_kernel<T><<<1, 256, 0, c.s>>>(nblk, pt + nblk, scal);
    cg_p_kernel<T><<<GL, 256, 0, c.s>>>(L, scal, R, P, ps.W, ps.wdof, ps.off, ps.Ps);
  }
  gcopy_kernel<T><<<GL, 256, 0, c.s>>>(L, X, R, g + G_RI_SKIP, nullptr);
  cg_end_kernel<<<1, 1, 0, c.s>>>(scal, g);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

// NSSOLVER (nssolver_dev) device-driven on Ri (4, N): the whole solve is
// enqueued on w.c.s with no host synchronisation.
template <class T>
int nssolver_graph(NsGraphWork<T> &w, T *Ri) {
  Ctx<T> &c = w.c;
  crimson_lhs *l = c.lhs;
  const int N = c.N, nsd = 3;
  const size_t mn = (size_t)nsd * N;
  double *g = w.gst.p;
  const double *ri = g + G_RI_SKIP;
  const unsigned gN = crimson::grid_for(N, 256), Gmn = crimson::grid_for(mn, 256);
  extract_kernel<<<gN, 256, 0, c.s>>>(N, 4, 0, 3, Ri, w.Rmi.p);
  extract_kernel<<<gN, 256, 0, c.s>>>(N, 4, 3, 1, Ri, w.Rci.p);
  gcopy_kernel<T><<<Gmn, 256, 0, c.s>>>(mn, w.Rmi.p, w.Rm.p, nullptr, nullptr);
  gcopy_kernel<T><<<gN, 256, 0, c.s>>>(N, w.Rci.p, w.Rc.p, nullptr, nullptr);
  CR_TRY(mdot<T>(c, mn, w.Rm.p, nullptr, 0, 0, true, c.dsc.p + 42, nullptr, true));
  CR_TRY(mdot<T>(c, N, w.Rc.p, nullptr, 0, 0, true, c.dsc.p + 43, nullptr, true));
  ns_begin_kernel<<<1, 1, 0, c.s>>>(g, c.dsc.p + 42, c.dsc.p + 43, w.riAbs, w.riRel);
  const T *V = planes(l, T(0));
  const MatView<T> mK = view<T>(c, V, 4, 0, 0, 3, 3, c.W.p, c.fix.p, 4);
  const MatView<T> mG = view<T>(c, V, 4, 0, 3, 3, 1, c.W.p, nullptr, 4);
  const MatView<T> mD = view<T>(c, V, 4, 3, 0, 1, 3, c.W.p, nullptr, 4);
  const MatView<T> mL = view<T>(c, V, 4, 3, 3, 1, 1, c.W.p, c.fix.p, 4);
  MatView<T> Gt = mD;
  if (!l->gt_is_d) {
    crimson::DBuf<T> &gtb = gt_buf(l, T(0));
    gt_kernel<T><<<crimson::grid_for(l->nnz, 256), 256, 0, c.s>>>(l->nnz, V, l->d_tr.p, gtb.p);
    Gt = view<T>(c, gtb.p, 1, 0, 0, 1, 3, c.W.p, nullptr, 4);
    Gt.plane[0] = 0;
    Gt.plane[1] = 1;
    Gt.plane[2] = 2;
    Gt.r0 = 3;
    Gt.c0 = 0;
  }
  // BCPRE and the faces' PRE coefficients on the device; the face
  // functionals (c.hfdev / c.fdev: set up once by ns_graph_setup)
  for (size_t f = 0; f < l->face.size(); ++f) {
    CrFace &fa = l->face[f];
    if (!fa.coupled) continue;
    face_pre_coef_kernel<<<1, 1, 0, c.s>>>(fa.nNo, fa.dof, nsd, fa.d_valM.p, fa.res, w.fcoef.p + f);
    if (fa.nsup > 0)
      face_q_kernel<T><<<crimson::grid_for(fa.nsup, 256), 256, 0, c.s>>>(
          fa.nsup, fa.d_qptr.p, fa.d_qa.p, fa.d_qk.p, fa.d_glob.p, fa.d_valM.p, fa.dof, mG, fa.d_q.p);
  }
  CR_HIP(hipGetLastError());
  const int nfc = (int)c.hfdev.size();
  T *t1 = w.t1.p, *t2 = w.t2.p, *tmp = w.gmTmp.p;
  auto schur_op = [&](const T *Pv, T *SPv, const double *scal) -> int {
    const double *done = scal + 5;
    const unsigned gr = crimson::grid_for((size_t)N * kNarrowLanes, 256);
    if (l->d_col16.p)
      g_product_kernel<T, short><<<gr + nfc * kFaceDotBlocks, 256, 0, c.s>>>(
          N, l->d_rowptr.p, l->d_col16.p, mG, t2, t1, tmp, nfc, c.fdev.p, c.fpart.p, done);
    else
      g_product_kernel<T, int><<<gr + nfc * kFaceDotBlocks, 256, 0, c.s>>>(
          N, l->d_rowptr.p, l->d_col.p, mG, t2, t1, tmp, nfc, c.fdev.p, c.fpart.p, done);
    if (w.disjoint && nfc > 1)
      face_pre_apply_multi_kernel<T><<<dim3(crimson::grid_for(std::max(1, w.maxfn), 256), nfc), 256, 0, c.s>>>(
          c.fdev.p, c.fpart.p, t1, tmp, c.W.p, 4, done);
    else
      for (int f = 0; f < nfc; ++f)
        if (c.hfdev[f].nNo > 0)
          face_pre_apply_kernel<T><<<crimson::grid_for(c.hfdev[f].nNo, 256), 256, 0, c.s>>>(
              c.hfdev[f], c.fpart.p + f * kFaceDotBlocks, t1, tmp, c.W.p, 4, done);
    if (l->d_col16.p)
      schur_kernel<T, short><<<gr, 256, 0, c.s>>>(N, l->d_rowptr.p, l->d_col16.p, mL, Gt, Pv, t2, tmp, SPv, done);
    else
      schur_kernel<T, int><<<gr, 256, 0, c.s>>>(N, l->d_rowptr.p, l->d_col.p, mL, Gt, Pv, t2, tmp, SPv, done);
    CR_HIP(hipGetLastError());
    return CRIMSON_OK;
  };
  Prescale<T> pre_p;
  pre_p.W = c.W.p;
  pre_p.wdof = 4;
  pre_p.off = mG.c0;
  pre_p.Ps = t2;
  const int nB = w.nB;
  for (int i = 1; i <= w.mRI; ++i) {
    const int iB = 2 * i - 1, iBB = 2 * i;
    T *Ui = w.U.p + mn * (i - 1), *Pi = w.P.p + (size_t)N * (i - 1);
    T *MUb = w.MU.p + mn * (iB - 1), *MUbb = w.MU.p + mn * (iBB - 1);
    T *MPb = w.MP.p + (size_t)N * (iB - 1), *MPbb = w.MP.p + (size_t)N * (iBB - 1);
    CR_TRY(w.cut());
    CR_TRY(gmres_graph(w, mK, 3, w.Rm.p, Ui));
    CR_TRY(w.cut());
    CR_TRY(spmv(c, mD, 1, 3, Ui, Pi, -1.0, w.Rc.p, 1.0, ri));  // P = Rc - D U
    CR_TRY(cg_graph(w, (size_t)N, Pi, schur_op, w.t3.p, w.t4.p, MPb, pre_p));
    CR_TRY(w.cut());
    CR_TRY(spmv(c, mG, 3, 1, Pi, MUb, 1.0, static_cast<const T *>(nullptr), 0.0, ri));
    gaxpby_kernel<T><<<Gmn, 256, 0, c.s>>>(mn, T(1), w.Rm.p, T(0), MUbb, ri);
    gaxpby_kernel<T><<<Gmn, 256, 0, c.s>>>(mn, T(-1), MUb, T(1), MUbb, ri);  // MU(iBB) = Rm - MU(iB)
    CR_TRY(gmres_graph(w, mK, 3, MUbb, Ui));
    CR_TRY(w.cut());
    CR_TRY(spmv(c, mK, 3, 3, Ui, MUbb, 1.0, static_cast<const T *>(nullptr), 0.0, ri));
    CR_TRY(addbcmul_graph(w, false, 3, Ui, MUbb, ri));
    CR_TRY(spmv(c, mL, 1, 1, Pi, MPb, 1.0, static_cast<const T *>(nullptr), 0.0, ri));
    CR_TRY(spmv(c, mD, 1, 3, Ui, MPbb, 1.0, static_cast<const T *>(nullptr), 0.0, ri));
    for (int k = iB; k <= iBB; ++k) {
      const T *MUk = w.MU.p + mn * (k - 1), *MPk = w.MP.p + (size_t)N * (k - 1);
      double *d = w.dots.p + (size_t)1024 * (k - iB);
      CR_TRY(mdot<T>(c, mn, MUk, w.MU.p, mn, k - 1, true, d, nullptr, true, ri));
      CR_TRY(mdot<T>(c, mn, MUk, w.Rmi.p, 0, 1, false, d + 256, nullptr, true, ri));
      CR_TRY(mdot<T>(c, N, MPk, w.MP.p, N, k - 1, true, d + 512, nullptr, true, ri));
      CR_TRY(mdot<T>(c, N, MPk, w.Rci.p, 0, 1, false, d + 768, nullptr, true, ri));
    }
    ns_gram_kernel<<<1, 1, 0, c.s>>>(g, i, nB, w.dots.p, w.G(w.oA), w.G(w.oBv), w.G(w.oxB), w.G(w.oC));
    gcopy_kernel<T><<<Gmn, 256, 0, c.s>>>(mn, w.Rmi.p, w.Rm.p, ri, nullptr);
    gcopy_kernel<T><<<gN, 256, 0, c.s>>>(N, w.Rci.p, w.Rc.p, ri, nullptr);
    gmaxpy_kernel<T><<<Gmn, 256, 0, c.s>>>(mn, w.Rm.p, w.MU.p, mn, iBB, nullptr, w.G(w.oxB), -1.0, ri, nullptr);
    gmaxpy_kernel<T><<<gN, 256, 0, c.s>>>(N, w.Rc.p, w.MP.p, N, iBB, nullptr, w.G(w.oxB), -1.0, ri, nullptr);
  }
  ns_final_kernel<<<1, 1, 0, c.s>>>(g, w.mRI, w.G(w.oxB), w.G(w.ocu), w.G(w.ocp));
  const double *err = g + G_ERR, *suc = g + G_RI_SUC;
  gcopy_kernel<T><<<gN, 256, 0, c.s>>>(N, w.Rci.p, w.Rc.p, err, suc);
  gmaxpy_kernel<T><<<gN, 256, 0, c.s>>>(N, w.Rc.p, w.MP.p, N, 0, g + G_NS_CNT, w.G(w.oxB), -1.0, err, suc);
  CR_TRY(mdot<T>(c, N, w.Rc.p, nullptr, 0, 0, true, g + G_RC_N2, nullptr, true, err));
  // solution: Rmi = sum cu(k) U_k, Rci = sum cp(k) P_k
  gfill_kernel<T><<<Gmn, 256, 0, c.s>>>(mn, w.Rmi.p, err);
  gfill_kernel<T><<<gN, 256, 0, c.s>>>(N, w.Rci.p, err);
  gmaxpy_kernel<T><<<Gmn, 256, 0, c.s>>>(mn, w.Rmi.p, w.U.p, mn, 0, g + G_RI_ITR, w.G(w.ocu), 1.0, err, nullptr);
  gmaxpy_kernel<T><<<gN, 256, 0, c.s>>>(N, w.Rci.p, w.P.p, N, 0, g + G_RI_ITR, w.G(w.ocp), 1.0, err, nullptr);
  insert_kernel<<<gN, 256, 0, c.s>>>(N, 4, 0, 3, w.Rmi.p, Ri);
  insert_kernel<<<gN, 256, 0, c.s>>>(N, 4, 3, 1, w.Rci.p, Ri);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

// The device-driven solve's workspaces and face set-up for lhs and ls
// (allocations, face supports and their uploads: all before any capture).
// res: memLS's face coefficients (res[f] != 0 couples Neumann face f), incL
// as in memLS_SOLVE.
template <class T>
int ns_graph_setup(NsGraphWork<T> &w, crimson_lhs *l, crimson_ls *ls, const int *incL, const double *res,
                   hipStream_t s) {
  if (ls->LS_type != CRIMSON_LS_NS) {
    crimson::set_error("the device-driven solve is the NS solver's (LS_TYPE_NS)");
    return CRIMSON_EUNSUPPORTED;
  }
  Ctx<T> &c = w.c;
  CR_TRY(c.init(l, 4));
  c.s = s;
  c.coupled = false;
  for (int f = 0; f < l->nFaces; ++f) {
    CrFace &fa = l->face[f];
    fa.inc = !(incL && incL[f] == 0);
    fa.coupled = false;
    if (!fa.inc || !fa.created) continue;
    if (fa.bGrp == CRIMSON_BC_NEU && res && res[f] != 0.0) {
      fa.res = res[f];
      fa.coupled = true;
      c.coupled = true;
    }
  }
  const int N = l->nNo;
  const size_t mn = (size_t)3 * N;
  w.mRI = ls->RI.mItr;
  w.mGM = ls->GM.mItr;
  w.mCG = ls->CG.mItr;
  w.sD = ls->GM.sD;
  w.nB = 2 * w.mRI;
  w.riAbs = ls->RI.absTol;
  w.riRel = ls->RI.relTol;
  w.gmAbs = ls->GM.absTol;
  w.gmRel = ls->GM.relTol;
  w.cgAbs = ls->CG.absTol;
  w.cgRel = ls->CG.relTol;
  if (w.mRI < 1 || w.mGM < 1 || w.mCG < 1 || w.sD < 1 || w.nB > 255 || w.sD > 255) {
    crimson::set_error("device-driven NS solve: iteration limits out of range (RI %d, GM %d x %d, CG %d)", w.mRI,
                       w.mGM, w.sD, w.mCG);
    return CRIMSON_EARG;
  }
  CR_TRY(w.Rm.alloc(mn));
  CR_TRY(w.Rc.alloc(N));
  CR_TRY(w.Rmi.alloc(mn));
  CR_TRY(w.Rci.alloc(N));
  CR_TRY(w.U.alloc(mn * w.mRI));
  CR_TRY(w.P.alloc((size_t)N * w.mRI));
  CR_TRY(w.MU.alloc(mn * w.nB));
  CR_TRY(w.MP.alloc((size_t)N * w.nB));
  CR_TRY(w.t1.alloc(mn));
  CR_TRY(w.t2.alloc(N));
  CR_TRY(w.t3.alloc(N));
  CR_TRY(w.t4.alloc(N));
  CR_TRY(w.gmTmp.alloc(std::max(mn, (size_t)4 * N)));  // GMRES scratch; the Schur CG's Z(4, N)
  CR_TRY(w.gmU.alloc(mn * (w.sD + 1)));
  if (sizeof(T) != sizeof(double)) CR_TRY(w.Rw.alloc((size_t)4 * N));
  size_t o = G_ARR;
  w.oH = o;
  o += (size_t)(w.sD + 1) * w.sD;
  w.ocs = o;
  o += w.sD;
  w.osn = o;
  o += w.sD;
  w.oerr = o;
  o += w.sD + 1;
  w.oy = o;
  o += w.sD;
  w.oA = o;
  o += (size_t)w.nB * w.nB;
  w.oBv = o;
  o += w.nB;
  w.oxB = o;
  o += w.nB;
  w.oC = o;
  o += (size_t)w.nB * (w.nB + 1);
  w.ocu = o;
  o += w.mRI;
  w.ocp = o;
  o += w.mRI;
  CR_TRY(w.gst.alloc(o));
  CR_TRY(w.fcoef.alloc(std::max<size_t>(1, l->face.size())));
  CR_TRY(w.dots.alloc(2048));
  if (!l->gt_is_d) CR_TRY(gt_buf(l, T(0)).alloc((size_t)3 * l->nnz));
  c.hfdev.clear();
  for (size_t f = 0; f < l->face.size(); ++f) {
    CrFace &fa = l->face[f];
    if (!fa.coupled) continue;
    CR_TRY(fa.d_valM.alloc((size_t)fa.dof * fa.nNo));
    CR_TRY(face_support(l, fa));
    c.hfdev.push_back(FaceDev{fa.d_glob.p, fa.d_valM.p, 0.0, w.fcoef.p + f, fa.res, fa.nNo, fa.dof, fa.d_sup.p,
                              fa.d_q.p, fa.nsup});
  }
  if (!c.hfdev.empty()) {
    CR_TRY(c.fdev.alloc(c.hfdev.size()));
    CR_HIP(hipMemcpyAsync(c.fdev.p, c.hfdev.data(), c.hfdev.size() * sizeof(FaceDev), hipMemcpyHostToDevice, s));
  }
  CR_TRY(w.msum.alloc(std::max<size_t>(1, c.hfdev.size())));
  {
    std::vector<char> seen(N, 0);
    w.disjoint = true;
    w.maxfn = 0;
    for (auto &fa : l->face) {
      if (!fa.coupled) continue;
      w.maxfn = std::max(w.maxfn, fa.nNo);
      for (int a = 0; a < fa.nNo; ++a) {
        if (seen[fa.glob[a]]) w.disjoint = false;
        seen[fa.glob[a]] = 1;
      }
    }
  }
  CR_HIP(hipStreamSynchronize(s));
  return CRIMSON_OK;
}

// One memLS NS solve (solve_impl) enqueued device-driven on w.c.s: PRECOND,
// R = W R, the solve, R W; d_Ri the FP64 (4, N) right-hand side in and
// solution out.  No synchronisation, no allocation.
template <class T>
int ns_graph_enqueue(NsGraphWork<T> &w, double *d_Ri) {
  Ctx<T> &c = w.c;
  crimson_lhs *l = c.lhs;
  const int N = l->nNo, dof = 4;
  const size_t L = (size_t)dof * N;
  const unsigned gN = crimson::grid_for(N, 256);
  const T *V = planes(l, T(0));
  precond_diag_kernel<T><<<gN, 256, 0, c.s>>>(N, dof, l->vdof, V, l->nnz, l->d_diag.p, c.W.p, c.fix.p);
  for (auto &fa : l->face)
    if (fa.created && fa.inc && fa.bGrp == CRIMSON_BC_DIR && fa.nNo > 0)
      precond_face_kernel<T><<<crimson::grid_for(fa.nNo, 256), 256, 0, c.s>>>(fa.nNo, dof, fa.dof, fa.d_glob.p,
                                                                              fa.d_val.p, c.W.p);
  precond_fix_kernel<T><<<gN, 256, 0, c.s>>>(N, dof, l->vdof, V, l->nnz, l->d_diag.p, c.W.p, c.fix.p);
  for (auto &fa : l->face)
    if (fa.coupled && fa.nNo > 0)
      face_valm_kernel<T><<<crimson::grid_for(fa.nNo, 256), 256, 0, c.s>>>(fa.nNo, dof, fa.dof, fa.d_glob.p,
                                                                            fa.d_val.p, c.W.p, fa.d_valM.p);
  T *R = sizeof(T) == sizeof(double) ? reinterpret_cast<T *>(d_Ri) : w.Rw.p;
  mul_kernel<double, T, T><<<crimson::grid_for(L, 256), 256, 0, c.s>>>(L, d_Ri, c.W.p, R);
  CR_HIP(hipGetLastError());
  CR_TRY(nssolver_graph<T>(w, R));
  mul_kernel<T, T, double><<<crimson::grid_for(L, 256), 256, 0, c.s>>>(L, R, c.W.p, d_Ri);  // R = R W
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

// The solve's statistics from a copy of the first G_ARR entries of gst
// into ls's fields (as nssolver_dev leaves them); CRIMSON_ESINGULAR when GE
// met a singular pivot.
int ns_graph_stats(const double *h, crimson_ls *ls) {
  if (h[G_ERR] != 0.0) {
    crimson::set_error("memLS GE: singular matrix");
    return CRIMSON_ESINGULAR;
  }
  ls->RI.iNorm = h[G_RI_INORM];
  ls->RI.fNorm = h[G_RI_FNORM];
  ls->RI.itr = (int)h[G_RI_ITR];
  ls->RI.suc = h[G_RI_SUC] != 0.0;
  ls->RI.dB = 10.0 * std::log(h[G_RI_FNORM] / h[G_RI_DBPRE]);
  ls->RI.callD = 0.0;
  ls->GM.iNorm = h[G_GM_INORM];
  ls->GM.fNorm = h[G_GM_FNORM];
  ls->GM.itr = (int)h[G_GM_ITR];
  ls->GM.suc = h[G_GM_SUC] != 0.0;
  ls->GM.dB = h[G_GM_LASTD] != 0.0 ? 10.0 * std::log(h[G_GM_LASTF] / h[G_GM_LASTD]) : 0.0;
  ls->GM.callD = 0.0;
  ls->CG.iNorm = h[G_CG_INORM];
  ls->CG.fNorm = std::sqrt(h[G_CG_ERR]);
  ls->CG.itr = (int)h[G_CG_ITR];
  ls->CG.suc = h[G_CG_SUC] != 0.0;
  ls->CG.dB = 5.0 * std::log(h[G_CG_ERR] / h[G_CG_ERRO]);
  ls->CG.callD = 0.0;
  const double rcn = std::sqrt(h[G_RC_N2]) / ls->RI.fNorm;
  ls->Resc = (int)std::lround(100.0 * rcn * rcn);
  ls->Resm = 100 - ls->Resc;
  return CRIMSON_OK;
}

// The signature of what a device-driven solve baked in: the matrix and
// face buffers it reads and the solver's settings.
std::vector<unsigned long long> ns_signature(const crimson_lhs *l, const crimson_ls *ls) {
  std::vector<unsigned long long> v = {l->d_V.serial, l->d_V32.serial, l->d_Gt.serial, l->d_Gt32.serial,
                                       (unsigned long long)l->vdof, (unsigned long long)l->prec,
                                       (unsigned long long)l->gt_is_d, (unsigned long long)ls->RI.mItr,
                                       (unsigned long long)ls->GM.mItr, (unsigned long long)ls->GM.sD,
                                       (unsigned long long)ls->CG.mItr};
  for (const auto &fa : l->face) {
    v.push_back(fa.d_glob.serial);
    v.push_back(fa.d_val.serial);
    v.push_back(fa.d_valM.serial);
    v.push_back(fa.d_sup.serial);
    v.push_back(fa.d_q.serial);
    v.push_back((unsigned long long)fa.coupled);
  }
  return v;
}

// A graph holding device-driven solves: the workspaces, and after a replay
// each captured solve's statistics (the last one's into ls).
template <class T>
struct NsGraphExtra : crimson::GraphExtra {
  NsGraphWork<T> w;
  crimson_lhs *lhs = nullptr;
  crimson_ls *ls = nullptr;
  int nsolve = 0;
  crimson::DBuf<double> stat;  // G_ARR doubles per captured solve
  std::vector<double> hstat;
  std::vector<unsigned long long> form_sig_unused;
  int after_launch() override {
    hstat.resize((size_t)G_ARR * nsolve);
    CR_HIP(hipMemcpy(hstat.data(), stat.p, hstat.size() * sizeof(double), hipMemcpyDeviceToHost));
    int rc = CRIMSON_OK;
    for (int k = 0; k < nsolve && rc == CRIMSON_OK; ++k) rc = ns_graph_stats(hstat.data() + (size_t)G_ARR * k, ls);
    return rc;
  }
  std::vector<unsigned long long> signature() const override { return ns_signature(lhs, ls); }
  int record_stats(int k) {
    CR_HIP(hipMemcpyAsync(stat.p + (size_t)G_ARR * k, w.gst.p, (size_t)G_ARR * sizeof(double),
                          hipMemcpyDeviceToDevice, w.c.s));
    return CRIMSON_OK;
  }
};

}  // namespace

extern "C" int crimson_ls_create(int lsType, int present, double relTol, double absTol, int maxItr, int dimKry,
                                 const double *relTolIn, const double *absTolIn, const int *maxItrIn,
                                 crimson_ls **out) {
  if (!out) {
    crimson::set_error("crimson_ls_create: null output");
    return CRIMSON_EARG;
  }
  crimson_ls *ls = new crimson_ls();
  ls->LS_type = lsType;
  switch (lsType) {  // memLS.f90:1394-1416
    case CRIMSON_LS_NS:
      ls->RI.relTol = 4e-1;
      ls->GM.relTol = 1e-2;
      ls->CG.relTol = 1e-1;
      ls->RI.mItr = 10;
      ls->GM.mItr = 3;
      ls->CG.mItr = 500;
      ls->GM.sD = 50;
      break;
    case CRIMSON_LS_GMRES:
      ls->RI.relTol = 1e-2;
      ls->RI.mItr = 2;
      ls->RI.sD = 150;
      break;
    case CRIMSON_LS_CG:
      ls->RI.relTol = 1e-4;
      ls->RI.mItr = 1000;
      break;
    default:
      delete ls;
      crimson::set_error("Solver type LS_TYPE %d is not defined", lsType);
      return CRIMSON_EARG;
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
  if ((present & 16) && relTolIn) {
    ls->GM.relTol = relTolIn[0];
    ls->CG.relTol = relTolIn[1];
  }
  if ((present & 32) && absTolIn) {
    ls->GM.absTol = absTolIn[0];
    ls->CG.absTol = absTolIn[1];
  }
  if ((present & 64) && maxItrIn) {
    ls->GM.mItr = maxItrIn[0];
    ls->CG.mItr = maxItrIn[1];
  }
  // NS keeps 2*maxItr least-squares columns, each a row of one projection
  // sweep (at most 255 per sweep); GMRES keeps dimKry+1 basis vectors.
  const bool bad_ns = lsType == CRIMSON_LS_NS && (ls->RI.mItr > 127 || ls->GM.sD < 1 || ls->GM.sD > 250);
  const bool bad_gm = lsType == CRIMSON_LS_GMRES && (ls->RI.sD < 1 || ls->RI.sD > 250);
  if (ls->RI.mItr < 1 || bad_ns || bad_gm) {
    crimson::set_error("solver limits out of range (maxItr %d, dimKry %d/%d; supported: maxItr >= 1, "
                       "NS maxItr <= 127, dimKry 1..250)",
                       ls->RI.mItr, ls->RI.sD, ls->GM.sD);
    delete ls;
    return CRIMSON_EUNSUPPORTED;
  }
  *out = ls;
  return CRIMSON_OK;
}

extern "C" int crimson_ls_free(crimson_ls *ls) {
  delete ls;
  return CRIMSON_OK;
}

extern "C" int crimson_lhs_create(int gnNo, int nNo, int nnz, const int *gNodes, const int *rowPtr,
                                  const int *colPtr, int nFaces, crimson_lhs **out) {
  if (!out || !rowPtr || !colPtr || nNo <= 0 || nnz <= 0 || nFaces < 0 || gnNo != nNo) {
    crimson::set_error("crimson_lhs_create: invalid arguments (single task: gnNo must equal nNo)");
    return CRIMSON_EARG;
  }
  if (gNodes)
    for (int a = 0; a < nNo; ++a)
      if (gNodes[a] != a + 1) {
        crimson::set_error("crimson_lhs_create: gNodes must be the identity on one task");
        return CRIMSON_EUNSUPPORTED;
      }
  crimson_lhs *l = new crimson_lhs();
  l->nNo = nNo;
  l->nnz = nnz;
  l->nFaces = nFaces;
  l->face.resize(nFaces);
  l->h_rowptr.resize(nNo + 1);
  l->h_col.resize(nnz);
  l->h_diag.assign(nNo, -1);
  l->h_tr.assign(nnz, -1);
  auto fail = [&](int code) {
    crimson_lhs_free(l);
    return code;
  };
  for (int i = 0; i <= nNo; ++i) l->h_rowptr[i] = rowPtr[i] - 1;
  if (l->h_rowptr[0] != 0 || l->h_rowptr[nNo] != nnz) {
    crimson::set_error("rowPtr must run from 1 to nnz+1");
    return fail(CRIMSON_EARG);
  }
  for (int k = 0; k < nnz; ++k) {
    l->h_col[k] = colPtr[k] - 1;
    if (l->h_col[k] < 0 || l->h_col[k] >= nNo) {
      crimson::set_error("colPtr(%d)=%d out of range", k + 1, colPtr[k]);
      return fail(CRIMSON_EARG);
    }
  }
  for (int i = 0; i < nNo; ++i) {
    if (l->h_rowptr[i + 1] < l->h_rowptr[i]) {
      crimson::set_error("rowPtr not monotone at %d", i + 1);
      return fail(CRIMSON_EARG);
    }
    for (int k = l->h_rowptr[i]; k < l->h_rowptr[i + 1]; ++k) {
      if (l->h_col[k] == i && l->h_diag[i] < 0) l->h_diag[i] = k;
      // transpose position (the DEPART search, memLS.f90:1668-1679, done once)
      const int j = l->h_col[k];
      for (int m = l->h_rowptr[j]; m < l->h_rowptr[j + 1]; ++m)
        if (l->h_col[m] == i) {
          l->h_tr[k] = m;
          break;
        }
    }
    if (l->h_diag[i] < 0) {
      crimson::set_error("row %d has no diagonal block", i + 1);
      return fail(CRIMSON_ENODIAG);
    }
  }
  for (int k = 0; k < nnz; ++k)
    if (l->h_tr[k] < 0) {
      crimson::set_error("CSR graph not symmetric at entry %d", k + 1);
      return fail(CRIMSON_EGRAPH);
    }
  if (hipStreamCreate(&l->stream) != hipSuccess) {
    l->stream = nullptr;
    crimson::set_error("hipStreamCreate failed (is a GPU visible?)");
    return fail(CRIMSON_EHIP);
  }
  if (hipEventCreate(&l->ev0) != hipSuccess || hipEventCreate(&l->ev1) != hipSuccess) {
    crimson::set_error("hipEventCreate failed");
    return fail(CRIMSON_EHIP);
  }
  int rc;
  if ((rc = l->d_rowptr.upload(l->h_rowptr.data(), nNo + 1, l->stream)) ||
      (rc = l->d_col.upload(l->h_col.data(), nnz, l->stream)) ||
      (rc = l->d_diag.upload(l->h_diag.data(), nNo, l->stream)) ||
      (rc = l->d_tr.upload(l->h_tr.data(), nnz, l->stream)))
    return fail(rc);
  {
    std::vector<short> c16(nnz);
    bool fits = true;
    for (int i = 0; i < nNo && fits; ++i)
      for (int k = l->h_rowptr[i]; k < l->h_rowptr[i + 1]; ++k) {
        const int d = l->h_col[k] - i;
        if (d < -32768 || d > 32767) {
          fits = false;
          break;
        }
        c16[k] = (short)d;
      }
    if (fits && (rc = l->d_col16.upload(c16.data(), nnz, l->stream))) return fail(rc);
  }
  if (hipStreamSynchronize(l->stream) != hipSuccess) {
    crimson::set_error("upload of the lhs structure failed");
    return fail(CRIMSON_EHIP);
  }
  *out = l;
  return CRIMSON_OK;
}

extern "C" int crimson_lhs_free(crimson_lhs *l) {
  if (!l) return CRIMSON_OK;
  if (l->stream) (void)hipStreamSynchronize(l->stream);
  if (l->ev0) (void)hipEventDestroy(l->ev0);
  if (l->ev1) (void)hipEventDestroy(l->ev1);
  hipStream_t s = l->stream;
  delete l;
  if (s) (void)hipStreamDestroy(s);
  return CRIMSON_OK;
}

extern "C" int crimson_bc_create(crimson_lhs *l, int faIn, int nNo, int dof, int bcType, const int *gNodes,
                                 const double *val) {
  if (!l || faIn < 1 || faIn > l->nFaces || nNo < 0 || dof < 1 || (nNo > 0 && !gNodes) ||
      (bcType != CRIMSON_BC_DIR && bcType != CRIMSON_BC_NEU)) {
    crimson::set_error("crimson_bc_create: invalid arguments (faIn=%d of %d)", faIn, l ? l->nFaces : 0);
    return CRIMSON_EARG;
  }
  CrFace &fa = l->face[faIn - 1];
  fa.nNo = nNo;
  fa.dof = dof;
  fa.qpat = false;
  fa.bGrp = bcType;
  fa.res = 0.0;
  fa.created = true;
  fa.glob.resize(nNo);
  for (int a = 0; a < nNo; ++a) {
    fa.glob[a] = gNodes[a] - 1;
    if (fa.glob[a] < 0 || fa.glob[a] >= l->nNo) {
      crimson::set_error("gNodes(%d)=%d out of range", a + 1, gNodes[a]);
      return CRIMSON_EARG;
    }
  }
  fa.val.assign((size_t)dof * nNo, 0.0);
  if (val) std::memcpy(fa.val.data(), val, sizeof(double) * dof * nNo);
  CR_TRY(fa.d_glob.upload(fa.glob.data(), nNo, l->stream));
  CR_TRY(fa.d_val.upload(fa.val.data(), fa.val.size(), l->stream));
  CR_TRY(fa.d_valM.alloc(fa.val.size()));
  CR_HIP(hipStreamSynchronize(l->stream));
  return CRIMSON_OK;
}

extern "C" int crimson_lhs_set_val(crimson_lhs *l, int dof, const double *Val) {
  if (!l || !Val || dof < 1 || dof > 4) {
    crimson::set_error("crimson_lhs_set_val: invalid arguments");
    return CRIMSON_EARG;
  }
  const size_t m = (size_t)dof * dof, nnz = l->nnz;
  crimson::DBuf<double> aos;
  CR_TRY(aos.upload(Val, m * nnz, l->stream));
  CR_TRY(l->d_V.alloc(m * nnz));
  aos_to_planes_kernel<double, double><<<crimson::grid_for(nnz, 256), 256, 0, l->stream>>>(nnz, (int)m, aos.p,
                                                                                            l->d_V.p);
  CR_HIP(hipGetLastError());
  CR_HIP(hipStreamSynchronize(l->stream));
  l->vdof = dof;
  l->prec = 64;
  l->gt_is_d = false;
  return CRIMSON_OK;
}

extern "C" int crimson_lhs_set_precision(crimson_lhs *l, int bits) {
  if (!l || (bits != 32 && bits != 64)) {
    crimson::set_error("crimson_lhs_set_precision: invalid arguments (bits 32 or 64)");
    return CRIMSON_EARG;
  }
  if (bits == l->prec) return CRIMSON_OK;
  const size_t n = (size_t)l->vdof * l->vdof * l->nnz;
  if (l->vdof == 0 || n == 0) {
    crimson::set_error("crimson_lhs_set_precision: no matrix held");
    return CRIMSON_ESTATE;
  }
  const unsigned G = crimson::grid_for(n, 256);
  if (bits == 32) {
    CR_TRY(l->d_V32.alloc(n));
    convert_kernel<double, float><<<G, 256, 0, l->stream>>>(n, l->d_V.p, l->d_V32.p);
  } else {
    CR_TRY(l->d_V.alloc(n));
    convert_kernel<float, double><<<G, 256, 0, l->stream>>>(n, l->d_V32.p, l->d_V.p);
  }
  CR_HIP(hipGetLastError());
  CR_HIP(hipStreamSynchronize(l->stream));
  l->prec = bits;
  return CRIMSON_OK;
}

template <class T>
__global__ void planes_to_aos_kernel(size_t nnz, int m, const T *__restrict__ V, double *__restrict__ aos) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  for (int p = 0; p < m; ++p) aos[(size_t)m * k + p] = (double)V[(size_t)p * nnz + k];
}

extern "C" int crimson_lhs_get_val(crimson_lhs *l, int dof, double *Val) {
  if (!l || !Val || dof != l->vdof || !(l->prec == 32 ? (const void *)l->d_V32.p : (const void *)l->d_V.p)) {
    crimson::set_error("crimson_lhs_get_val: no matrix of block size %d held", dof);
    return CRIMSON_EARG;
  }
  const size_t m = (size_t)dof * dof, nnz = l->nnz;
  crimson::DBuf<double> aos;
  CR_TRY(aos.alloc(m * nnz));
  if (l->prec == 32)
    planes_to_aos_kernel<float><<<crimson::grid_for(nnz, 256), 256, 0, l->stream>>>(nnz, (int)m, l->d_V32.p, aos.p);
  else
    planes_to_aos_kernel<double><<<crimson::grid_for(nnz, 256), 256, 0, l->stream>>>(nnz, (int)m, l->d_V.p, aos.p);
  CR_HIP(hipGetLastError());
  CR_HIP(hipMemcpyAsync(Val, aos.p, m * nnz * sizeof(double), hipMemcpyDeviceToHost, l->stream));
  CR_HIP(hipStreamSynchronize(l->stream));
  return CRIMSON_OK;
}

// Block transpose: block k = (i, j) of dst is block (j, i) of src transposed,
// B'_k(r, c) = B_tr(k)(c, r), every plane of every block.
template <class T>
__global__ void transpose_blocks_kernel(size_t nnz, int m, const int *__restrict__ tr, const T *__restrict__ V,
                                        T *__restrict__ Vt) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const size_t t = (size_t)tr[k];
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) Vt[(size_t)(r * m + c) * nnz + k] = V[(size_t)(c * m + r) * nnz + t];
}

extern "C" int crimson_lhs_transpose_dev(crimson_lhs *dst, const crimson_lhs *src) {
  if (!dst || !src || dst == src) {
    crimson::set_error("crimson_lhs_transpose_dev: two distinct handles needed");
    return CRIMSON_EARG;
  }
  if (dst->nNo != src->nNo || dst->nnz != src->nnz || dst->h_rowptr != src->h_rowptr || dst->h_col != src->h_col) {
    crimson::set_error("crimson_lhs_transpose_dev: the handles' CSR graphs differ");
    return CRIMSON_EARG;
  }
  const bool f32 = src->prec == 32;
  if (src->vdof < 1 || !(f32 ? (const void *)src->d_V32.p : (const void *)src->d_V.p)) {
    crimson::set_error("crimson_lhs_transpose_dev: src holds no matrix");
    return CRIMSON_ESTATE;
  }
  const int m = src->vdof;
  const size_t nnz = src->nnz, n = (size_t)m * m * nnz;
  CR_HIP(hipStreamSynchronize(src->stream));
  const unsigned G = crimson::grid_for(nnz, 256);
  if (f32) {
    CR_TRY(dst->d_V32.alloc(n));
    transpose_blocks_kernel<float><<<G, 256, 0, dst->stream>>>(nnz, m, dst->d_tr.p, src->d_V32.p, dst->d_V32.p);
  } else {
    CR_TRY(dst->d_V.alloc(n));
    transpose_blocks_kernel<double><<<G, 256, 0, dst->stream>>>(nnz, m, dst->d_tr.p, src->d_V.p, dst->d_V.p);
  }
  CR_HIP(hipGetLastError());
  CR_HIP(hipStreamSynchronize(dst->stream));
  dst->vdof = m;
  dst->prec = src->prec;
  // a formation-produced matrix has D = -G^T blockwise; its transpose keeps
  // that relation (D'_ij = G_ji^T = -D_ij = -G'_ji^T)
  dst->gt_is_d = src->gt_is_d;
  return CRIMSON_OK;
}

extern "C" int crimson_ls_solve_dev(crimson_lhs *l, crimson_ls *ls, int dof, double *d_Ri, const int *incL,
                                    const double *res) {
  if (!l || !ls || !d_Ri) {
    crimson::set_error("crimson_ls_solve_dev: null argument");
    return CRIMSON_EARG;
  }
  return solve_dev(l, ls, dof, d_Ri, incL, res);
}

extern "C" int crimson_ls_solve(crimson_lhs *l, crimson_ls *ls, int dof, double *Ri, const double *Val,
                                const int *incL, const double *res) {
  if (!l || !ls || !Ri || !Val) {
    crimson::set_error("crimson_ls_solve: null argument");
    return CRIMSON_EARG;
  }
  CR_TRY(crimson_lhs_set_val(l, dof, Val));
  crimson::DBuf<double> dR;
  const size_t L = (size_t)dof * l->nNo;
  CR_TRY(dR.upload(Ri, L, l->stream));
  CR_TRY(solve_dev(l, ls, dof, dR.p, incL, res));
  CR_HIP(hipMemcpyAsync(Ri, dR.p, L * sizeof(double), hipMemcpyDeviceToHost, l->stream));
  CR_HIP(hipStreamSynchronize(l->stream));
  return CRIMSON_OK;
}

extern "C" int crimson_ls_stats(const crimson_ls *ls, int sub, double *o) {
  if (!ls || !o || sub < 0 || sub > 2) {
    crimson::set_error("crimson_ls_stats: invalid arguments");
    return CRIMSON_EARG;
  }
  const SubLs &s = sub == 0 ? ls->RI : (sub == 1 ? ls->GM : ls->CG);
  o[0] = s.iNorm;
  o[1] = s.fNorm;
  o[2] = s.itr;
  o[3] = s.suc;
  o[4] = s.dB;
  o[5] = s.callD;
  o[6] = ls->Resm;
  o[7] = ls->Resc;
  return CRIMSON_OK;
}

namespace {
template <class T>
int spmv_timed(crimson_lhs *l, int dof, const T *x, T *y, int reps, double *ms) {
  Ctx<T> c;
  CR_TRY(c.init(l, 4));
  const unsigned gN = crimson::grid_for(l->nNo, 256);
  const T *V = planes(l, T(0));
  precond_diag_kernel<T><<<gN, 256, 0, c.s>>>(l->nNo, 4, 4, V, l->nnz, l->d_diag.p, c.W.p, c.fix.p);
  const MatView<T> A = view<T>(c, V, 4, 0, 0, dof, dof, c.W.p, c.fix.p, 4);
  CR_TRY(spmv<T>(c, A, dof, dof, x, y));  // warm-up
  CR_HIP(hipEventRecord(l->ev0, c.s));
  for (int r = 0; r < reps; ++r) CR_TRY(spmv<T>(c, A, dof, dof, x, y));
  CR_HIP(hipEventRecord(l->ev1, c.s));
  CR_HIP(hipEventSynchronize(l->ev1));
  float t = 0.f;
  CR_HIP(hipEventElapsedTime(&t, l->ev0, l->ev1));
  *ms = t / reps;
  return CRIMSON_OK;
}
}  // namespace

extern "C" int crimson_lhs_spmv_dev(crimson_lhs *l, int dof, const double *d_x, double *d_y, int reps, double *ms) {
  if (!l || !d_x || !d_y || reps < 1 || !ms || (dof != 3 && dof != 4) || l->vdof != 4 ||
      !(l->prec == 32 ? (const void *)l->d_V32.p : (const void *)l->d_V.p)) {
    crimson::set_error("crimson_lhs_spmv_dev: invalid arguments (needs a 4x4 matrix, dof 3 or 4)");
    return CRIMSON_EARG;
  }
  if (l->prec == 64) return spmv_timed<double>(l, dof, d_x, d_y, reps, ms);
  // FP32 matrix: the product runs on FP32 copies of x and y
  const size_t L = (size_t)dof * l->nNo;
  crimson::DBuf<float> xf, yf;
  CR_TRY(xf.alloc(L));
  CR_TRY(yf.alloc(L));
  convert_kernel<double, float><<<crimson::grid_for(L, 256), 256, 0, l->stream>>>(L, d_x, xf.p);
  CR_HIP(hipGetLastError());
  CR_TRY(spmv_timed<float>(l, dof, xf.p, yf.p, reps, ms));
  convert_kernel<float, double><<<crimson::grid_for(L, 256), 256, 0, l->stream>>>(L, yf.p, d_y);
  CR_HIP(hipGetLastError());
  CR_HIP(hipStreamSynchronize(l->stream));
  return CRIMSON_OK;
}

namespace {
// Launches per captured segment at most: a device-driven solve enqueues
// every possible iteration, so its launch count grows with its iteration
// limits; the capture is cut into segments (each GMRES call, each CG call,
// the glue between them) and a configuration whose segment would exceed
// this is refused before it is captured.
constexpr size_t kMaxSegmentLaunches = 40000;

// A stream capture cut into segments (crimson_graph::seg), each
// instantiated on its own and launched back to back.
struct SegCapture {
  crimson_graph *g;
  hipStream_t s;
  bool open = false;
  int begin() {
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      crimson::set_error("hipStreamBeginCapture failed");
      return CRIMSON_EHIP;
    }
    open = true;
    return CRIMSON_OK;
  }
  int close() {
    hipGraph_t gr = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &gr);
    open = false;
    if (ec != hipSuccess || !gr) {
      crimson::set_error("capture into a HIP graph failed");
      return CRIMSON_EHIP;
    }
    size_t n = 0;
    if (hipGraphGetNodes(gr, nullptr, &n) != hipSuccess) {
      (void)hipGraphDestroy(gr);
      crimson::set_error("hipGraphGetNodes failed");
      return CRIMSON_EHIP;
    }
    if (n == 0) {
      (void)hipGraphDestroy(gr);
      return CRIMSON_OK;
    }
    g->seg.push_back(gr);
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, gr, nullptr, nullptr, 0) != hipSuccess) {
      crimson::set_error("hipGraphInstantiate failed");
      return CRIMSON_EHIP;
    }
    g->segx.push_back(x);
    return CRIMSON_OK;
  }
  int cut() {
    CR_TRY(close());
    return begin();
  }
  void abort() {
    if (!open) return;
    hipGraph_t gr = nullptr;
    (void)hipStreamEndCapture(s, &gr);
    if (gr) (void)hipGraphDestroy(gr);
    open = false;
  }
};

// The launches one segment of a device-driven solve holds at most, from the
// solver's limits (GMRES: per Arnoldi step about 16 launches, every coupled
// face batched or not; CG: about 8 per iteration), checked before capturing.
int check_segment_budget(int mGM, int sD, int mCG, int nfc_launch) {
  const size_t gm = (size_t)mGM * ((size_t)sD * (16 + 4 * nfc_launch) + 16);
  const size_t cg = (size_t)mCG * (8 + nfc_launch) + 8;
  if (gm > kMaxSegmentLaunches || cg > kMaxSegmentLaunches) {
    crimson::set_error("device-driven NS solve: a GMRES call of %d cycles x %d steps (%zu launches) or a CG call of "
                       "%d iterations (%zu launches) exceeds the %zu launches a captured segment may hold",
                       mGM, sD, gm, mCG, cg, kMaxSegmentLaunches);
    return CRIMSON_EUNSUPPORTED;
  }
  return CRIMSON_OK;
}

template <class T>
int NsGraphWork<T>::cut() {
  return cap ? cap->cut() : CRIMSON_OK;
}

template <class T>
int capture_into(crimson_graph *g, hipStream_t s, NsGraphWork<T> &w, const std::function<int()> &body) {
  SegCapture cap{g, s};
  const int nfc = (int)w.c.hfdev.size();
  CR_TRY(check_segment_budget(w.mGM, w.sD, w.mCG, (w.disjoint && nfc > 1) ? 1 : nfc));
  CR_TRY(cap.begin());
  w.cap = &cap;
  int rc = body();
  if (!rc) rc = cap.close();
  w.cap = nullptr;
  if (rc) cap.abort();
  g->s = s;
  return rc;
}

void graph_discard(crimson_graph *g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->g) (void)hipGraphDestroy(g->g);
  for (hipGraphExec_t x : g->segx) (void)hipGraphExecDestroy(x);
  for (hipGraph_t h : g->seg) (void)hipGraphDestroy(h);
  delete g->extra;
  delete g;
}

template <class T>
int ls_graph_create_t(crimson_lhs *l, crimson_ls *ls, double *d_Ri, const int *incL, const double *res,
                      crimson_graph **out) {
  crimson_graph *g = new crimson_graph();
  auto *ex = new NsGraphExtra<T>();
  g->extra = ex;
  ex->lhs = l;
  ex->ls = ls;
  ex->nsolve = 1;
  int rc = ns_graph_setup<T>(ex->w, l, ls, incL, res, l->stream);
  if (!rc) rc = ex->stat.alloc(G_ARR);
  if (!rc)
    rc = capture_into<T>(g, l->stream, ex->w, [&]() -> int {
      CR_TRY(ns_graph_enqueue<T>(ex->w, d_Ri));
      return ex->record_stats(0);
    });
  if (rc) {
    graph_discard(g);
    return rc;
  }
  g->lhs = l;
  g->sig = ex->signature();
  *out = g;
  return CRIMSON_OK;
}

template <class T>
int step_graph_create_t(crimson_form *f, crimson_lhs *l, crimson_ls *ls, double *d_yold, double *d_acold,
                        double *d_y, double *d_ac, double *d_ya, double *d_aca, const double *d_x, double *d_Res4,
                        const int *d_iBC, const double *d_BC, int ndofBC, int nsurf, const int *surf_ptr,
                        const int *d_glob, const double *d_nabi, const double *c1, const double *c2, double *d_q,
                        const double *faceRes, int nonlinear, crimson_graph **out) {
  crimson_graph *g = new crimson_graph();
  auto *ex = new NsGraphExtra<T>();
  g->extra = ex;
  ex->lhs = l;
  ex->ls = ls;
  ex->nsolve = nonlinear;
  hipStream_t s = crimson::form_stream(f);
  const int N = crimson::form_nshg(f);
  int rc = crimson::form_prepare_to_lhs(f, d_x, l);
  if (!rc) rc = ns_graph_setup<T>(ex->w, l, ls, nullptr, faceRes, s);
  if (!rc) rc = ex->stat.alloc((size_t)G_ARR * nonlinear);
  const crimson_form_params prm = crimson::form_params_of(f);
  if (!rc)
    rc = capture_into<T>(g, s, ex->w, [&]() -> int {
      CR_TRY(crimson_itr_predict_dev(f, d_yold, d_acold, d_y, d_ac, prm.gami));
      CR_TRY(crimson::form_wall_enqueue_predict(f, d_yold, d_acold));
      CR_TRY(crimson_itr_bc_dev(f, d_iBC, d_BC, ndofBC, d_y));
      for (int k = 0; k < nonlinear; ++k) {
        CR_TRY(crimson::form_wall_enqueue_alpha(f));
        CR_TRY(crimson_itr_yalpha_dev(f, d_yold, d_acold, d_y, d_ac, d_ya, d_aca, prm.alfi, prm.almi));
        CR_TRY(crimson::form_enqueue_to_lhs(f, d_ya, d_aca, d_x, d_Res4, l));
        // ElmpvsQ, then bc3Res (elmgmr.f90:528-555)
        for (int q = 0; q < nsurf; ++q)
          CR_TRY(crimson_elmpvsq_coef_dev(f, surf_ptr[q + 1] - surf_ptr[q], d_glob + surf_ptr[q],
                                          d_nabi + (size_t)3 * surf_ptr[q], c1[q], c2[q], d_ya, d_Res4, d_q + q));
        CR_TRY(crimson::form_enqueue_bc3res(f, d_Res4, 1));
        CR_TRY(ns_graph_enqueue<T>(ex->w, d_Res4));
        CR_TRY(ex->record_stats(k));
        CR_TRY(crimson_itr_correct_dev(f, d_y, d_ac, d_Res4, prm.gami, prm.alfi, prm.Delt));
        CR_TRY(crimson::form_wall_enqueue_correct(f, d_Res4));
        CR_TRY(crimson_itr_bc_dev(f, d_iBC, d_BC, ndofBC, d_y));
      }
      CR_HIP(hipMemcpyAsync(d_yold, d_y, (size_t)4 * N * sizeof(double), hipMemcpyDeviceToDevice, s));
      CR_HIP(hipMemcpyAsync(d_acold, d_ac, (size_t)4 * N * sizeof(double), hipMemcpyDeviceToDevice, s));
      CR_TRY(crimson::form_wall_enqueue_update(f));
      return CRIMSON_OK;
    });
  if (rc) {
    graph_discard(g);
    return rc;
  }
  g->f = f;
  g->lhs = l;
  g->sig = crimson::graph_signature(f, l);
  const std::vector<unsigned long long> e = ex->signature();
  g->sig.insert(g->sig.end(), e.begin(), e.end());
  *out = g;
  return CRIMSON_OK;
}
}  // namespace

extern "C" int crimson_ls_graph_create(crimson_lhs *lhs, crimson_ls *ls, int dof, double *d_Ri, const int *incL,
                                       const double *res, crimson_graph **out) {
  if (!lhs || !ls || !d_Ri || !out || dof != 4) {
    crimson::set_error("crimson_ls_graph_create: invalid arguments (the NS solver, dof 4)");
    return CRIMSON_EARG;
  }
  if (lhs->vdof != 4 || !(lhs->prec == 32 ? (const void *)lhs->d_V32.p : (const void *)lhs->d_V.p)) {
    crimson::set_error("crimson_ls_graph_create: lhs holds no 4x4 matrix");
    return CRIMSON_ESTATE;
  }
  if (lhs->prec == 32) return ls_graph_create_t<float>(lhs, ls, d_Ri, incL, res, out);
  return ls_graph_create_t<double>(lhs, ls, d_Ri, incL, res, out);
}

extern "C" int crimson_step_graph_create(crimson_form *f, crimson_lhs *lhs, crimson_ls *ls, double *d_yold,
                                         double *d_acold, double *d_y, double *d_ac, double *d_yAlpha,
                                         double *d_acAlpha, const double *d_x, double *d_Res4, const int *d_iBC,
                                         const double *d_BC, int ndofBC, int nsurf, const int *surf_ptr,
                                         const int *d_glob, const double *d_nabi, const double *c1,
                                         const double *c2, double *d_q, const double *faceRes, int nonlinear,
                                         crimson_graph **out) {
  if (!f || !lhs || !ls || !d_yold || !d_acold || !d_y || !d_ac || !d_yAlpha || !d_acAlpha || !d_x || !d_Res4 ||
      !d_iBC || !d_BC || ndofBC < 6 || nsurf < 0 || (nsurf > 0 && (!surf_ptr || !d_glob || !d_nabi || !c1 || !c2 ||
                                                                    !d_q)) ||
      nonlinear < 1 || !out) {
    crimson::set_error("crimson_step_graph_create: invalid arguments");
    return CRIMSON_EARG;
  }
  CR_TRY(crimson::form_wall_step_ready(f));
  for (int q = 0; q < nsurf; ++q)
    if (surf_ptr[q + 1] < surf_ptr[q]) {
      crimson::set_error("crimson_step_graph_create: surf_ptr not monotone");
      return CRIMSON_EARG;
    }
  if (crimson::form_params_of(f).precision == 32)
    return step_graph_create_t<float>(f, lhs, ls, d_yold, d_acold, d_y, d_ac, d_yAlpha, d_acAlpha, d_x, d_Res4,
                                      d_iBC, d_BC, ndofBC, nsurf, surf_ptr, d_glob, d_nabi, c1, c2, d_q, faceRes,
                                      nonlinear, out);
  return step_graph_create_t<double>(f, lhs, ls, d_yold, d_acold, d_y, d_ac, d_yAlpha, d_acAlpha, d_x, d_Res4,
                                     d_iBC, d_BC, ndofBC, nsurf, surf_ptr, d_glob, d_nabi, c1, c2, d_q, faceRes,
                                     nonlinear, out);
}
