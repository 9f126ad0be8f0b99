// Device kernels for the steps either side of the formation/solve path in one
// nonlinear iteration of itrdrv_iter_step (itrdrv.f90:1284-1528):
// itrPredict, itrYAlpha, itrCorrect (itrPC.f90), itrBC (itrbc.f90) for the
// supported BC codes, and the resistance-outlet residual term of ElmpvsQ
// (elmgmr.f90:1152-1176).  Fields are y(nshg,4), ac(nshg,4) column-major,
// the solver increment is Res4(4,nshg) as memLS returns it (solfar.f90:279).

#include <hip/hip_runtime.h>

#include "common.hpp"

using crimson::form_nshg;
using crimson::form_stream;

namespace {

// itrPredict, ipred = 1 (itrPC.f90:104-108): y = yold, ac = acold (gami-1)/gami
__global__ void predict_kernel(size_t n4, const double *__restrict__ yold, const double *__restrict__ acold,
                               double *__restrict__ y, double *__restrict__ ac, double fct) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n4) return;
  y[k] = yold[k];
  ac[k] = acold[k] * fct;
}

// itrYAlpha (itrPC.f90:326-334)
__global__ void yalpha_kernel(int N, const double *__restrict__ yold, const double *__restrict__ acold,
                              const double *__restrict__ y, const double *__restrict__ ac,
                              double *__restrict__ ya, double *__restrict__ aca, double alfi, double almi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const size_t n = N;
  aca[i + 3 * n] = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ya[i + c * n] = yold[i + c * n] + alfi * (y[i + c * n] - yold[i + c * n]);
    aca[i + c * n] = acold[i + c * n] + almi * (ac[i + c * n] - acold[i + c * n]);
  }
  ya[i + 3 * n] = y[i + 3 * n];
}

// itrCorrect (itrPC.f90:176-181), solinc(i,1:4) = Res4(1:4,i)
__global__ void correct_kernel(int N, double *__restrict__ y, double *__restrict__ ac,
                               const double *__restrict__ res4, double fct1, double fct2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const size_t n = N;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double d = res4[4 * (size_t)i + c];
    y[i + c * n] = y[i + c * n] + fct1 * d;
    ac[i + c * n] = ac[i + c * n] + d;
  }
  y[i + 3 * n] = y[i + 3 * n] + fct2 * res4[4 * (size_t)i + 3];
}

// itrBC (itrbc.f90:20-98) without periodic nodes: the constrained velocity
// components from BC(:,3:6) for codes 1-7, the pressure from BC(:,1) for
// bit 2.  BC(N, ndofBC), ndofBC >= 6.
__global__ void itrbc_kernel(int N, const int *__restrict__ ibc, const double *__restrict__ BC, int ndofBC,
                             double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int c = ibc[i];
  const size_t n = N;
  const int vb = (c >> 3) & 7;
  if (vb != 0) {
    const double b3 = BC[i + 2 * n], b4 = BC[i + 3 * n], b5 = BC[i + 4 * n], b6 = BC[i + 5 * n];
    const double u1 = y[i], u2 = y[i + n], u3 = y[i + 2 * n];
    switch (vb) {
      case 1:
        y[i] = b3 - b4 * u2 - b5 * u3;
        break;
      case 2:
        y[i + n] = b3 - b4 * u1 - b5 * u3;
        break;
      case 3:
        y[i] = b3 - b4 * u3;
        y[i + n] = b5 - b6 * u3;
        break;
      case 4:
        y[i + 2 * n] = b3 - b4 * u1 - b5 * u2;
        break;
      case 5:
        y[i] = b3 - b4 * u2;
        y[i + 2 * n] = b5 - b6 * u2;
        break;
      case 6:
        y[i + n] = b3 - b4 * u1;
        y[i + 2 * n] = b5 - b6 * u1;
        break;
      default:  // 7
        y[i] = b3;
        y[i + n] = b4;
        y[i + 2 * n] = b5;
        break;
    }
  }
  if ((c >> 2) & 1) y[i + 3 * n] = BC[i];
  (void)ndofBC;
}

// Q = sum_a NABI(a,:) . u(glob(a)) in one workgroup (GetFlowQ through NABI)
__global__ void __launch_bounds__(256) flowq_kernel(int nf, const int *__restrict__ glob,
                                                    const double *__restrict__ nabi, const double *__restrict__ y,
                                                    int N, double *__restrict__ q) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int a = threadIdx.x; a < nf; a += 256) {
    const int g = glob[a];
    s += nabi[3 * (size_t)a] * y[g] + nabi[3 * (size_t)a + 1] * y[g + (size_t)N] +
         nabi[3 * (size_t)a + 2] * y[g + 2 * (size_t)N];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) *q = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// res(i,1:3) += p NABI(i,1:3) with p = -(c1 Q + c2): the residual side of
// every coupled surface of ElmpvsQ (elmgmr.f90:1153-1330; a resistance has
// c1 = R, c2 = 0)
__global__ void pvsq_kernel(int nf, const int *__restrict__ glob, const double *__restrict__ nabi,
                            const double *__restrict__ q, double c1, double c2, double *__restrict__ res4) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nf) return;
  const double p = -(c1 * *q + c2);
  const size_t g = glob[a];
#pragma unroll
  for (int c = 0; c < 3; ++c) res4[4 * g + c] = res4[4 * g + c] + p * nabi[3 * (size_t)a + c];
}

// Adjoint of the coupled-surface term for a cotangent w4 laid out like
// Res4(4,nshg): s = sum_a w4(1:3, glob a) . NABI(a,:) and Q = sum NABI . u in
// one workgroup (fixed order), then cbar = (-Q s, -s), the cotangents of c1
// and c2.
__global__ void __launch_bounds__(256) pvsq_vjp_dot_kernel(int nf, const int *__restrict__ glob,
                                                           const double *__restrict__ nabi,
                                                           const double *__restrict__ y, int N,
                                                           const double *__restrict__ w4,
                                                           double *__restrict__ cbar) {
  __shared__ double sh[2][4];
  double s = 0.0, q = 0.0;
  for (int a = threadIdx.x; a < nf; a += 256) {
    const size_t g = glob[a];
    const double *nb = nabi + 3 * (size_t)a;
    s += nb[0] * w4[4 * g] + nb[1] * w4[4 * g + 1] + nb[2] * w4[4 * g + 2];
    q += nb[0] * y[g] + nb[1] * y[g + (size_t)N] + nb[2] * y[g + 2 * (size_t)N];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off, 64);
    q += __shfl_xor(q, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x / 64] = s;
    sh[1][threadIdx.x / 64] = q;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double ss = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    const double qq = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    cbar[0] = -qq * ss;
    cbar[1] = -ss;
  }
}

// ybar(glob a, 1:3) += c1 cbar[1] NABI(a,:), ybar (nshg,4) column-major
__global__ void pvsq_vjp_axpy_kernel(int nf, const int *__restrict__ glob, const double *__restrict__ nabi, int N,
                                     double c1, const double *__restrict__ cbar, double *__restrict__ ybar) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nf) return;
  const double f = c1 * cbar[1];
  const size_t g = glob[a];
#pragma unroll
  for (int c = 0; c < 3; ++c) ybar[g + c * (size_t)N] = ybar[g + c * (size_t)N] + f * nabi[3 * (size_t)a + c];
}

}  // namespace

extern "C" int crimson_itr_predict_dev(crimson_form *f, const double *d_yold, const double *d_acold, double *d_y,
                                       double *d_ac, double gami) {
  if (!f || !d_yold || !d_acold || !d_y || !d_ac || gami == 0.0) {
    crimson::set_error("crimson_itr_predict_dev: invalid arguments");
    return CRIMSON_EARG;
  }
  const size_t n4 = 4 * (size_t)form_nshg(f);
  predict_kernel<<<crimson::grid_for(n4, 256), 256, 0, form_stream(f)>>>(n4, d_yold, d_acold, d_y, d_ac,
                                                                                 (gami - 1.0) / gami);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

extern "C" int crimson_itr_yalpha_dev(crimson_form *f, const double *d_yold, const double *d_acold,
                                      const double *d_y, const double *d_ac, double *d_yAlpha, double *d_acAlpha,
                                      double alfi, double almi) {
  if (!f || !d_yold || !d_acold || !d_y || !d_ac || !d_yAlpha || !d_acAlpha) {
    crimson::set_error("crimson_itr_yalpha_dev: invalid arguments");
    return CRIMSON_EARG;
  }
  const int N = form_nshg(f);
  yalpha_kernel<<<crimson::grid_for(N, 256), 256, 0, form_stream(f)>>>(N, d_yold, d_acold, d_y, d_ac,
                                                                               d_yAlpha, d_acAlpha, alfi, almi);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

extern "C" int crimson_itr_correct_dev(crimson_form *f, double *d_y, double *d_ac, const double *d_Res4,
                                       double gami, double alfi, double Delt) {
  if (!f || !d_y || !d_ac || !d_Res4) {
    crimson::set_error("crimson_itr_correct_dev: invalid arguments");
    return CRIMSON_EARG;
  }
  const int N = form_nshg(f);
  correct_kernel<<<crimson::grid_for(N, 256), 256, 0, form_stream(f)>>>(N, d_y, d_ac, d_Res4, gami * Delt,
                                                                                gami * alfi * Delt);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

extern "C" int crimson_itr_bc_dev(crimson_form *f, const int *d_iBC, const double *d_BC, int ndofBC, double *d_y) {
  if (!f || !d_iBC || !d_BC || !d_y || ndofBC < 6) {
    crimson::set_error("crimson_itr_bc_dev: invalid arguments (BC(nshg,ndofBC) needs ndofBC >= 6)");
    return CRIMSON_EARG;
  }
  const int N = form_nshg(f);
  itrbc_kernel<<<crimson::grid_for(N, 256), 256, 0, form_stream(f)>>>(N, d_iBC, d_BC, ndofBC, d_y);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

extern "C" int crimson_elmpvsq_dev(crimson_form *f, int nface, const int *d_glob, const double *d_nabi, double R,
                                   const double *d_y, double *d_Res4, double *d_q) {
  if (!f || nface < 0 || (nface > 0 && (!d_glob || !d_nabi)) || !d_y || !d_Res4 || !d_q) {
    crimson::set_error("crimson_elmpvsq_dev: invalid arguments");
    return CRIMSON_EARG;
  }
  const int N = form_nshg(f);
  hipStream_t s = form_stream(f);
  flowq_kernel<<<1, 256, 0, s>>>(nface, d_glob, d_nabi, d_y, N, d_q);
  if (nface > 0)
    pvsq_kernel<<<crimson::grid_for(nface, 256), 256, 0, s>>>(nface, d_glob, d_nabi, d_q, R, 0.0, d_Res4);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

extern "C" int crimson_elmpvsq_coef_dev(crimson_form *f, int nface, const int *d_glob, const double *d_nabi,
                                        double c1, double c2, const double *d_y, double *d_Res4, double *d_q) {
  if (!f || nface < 0 || (nface > 0 && (!d_glob || !d_nabi)) || !d_y || !d_Res4 || !d_q) {
    crimson::set_error("crimson_elmpvsq_coef_dev: invalid arguments");
    return CRIMSON_EARG;
  }
  const int N = form_nshg(f);
  hipStream_t s = form_stream(f);
  flowq_kernel<<<1, 256, 0, s>>>(nface, d_glob, d_nabi, d_y, N, d_q);
  if (nface > 0)
    pvsq_kernel<<<crimson::grid_for(nface, 256), 256, 0, s>>>(nface, d_glob, d_nabi, d_q, c1, c2, d_Res4);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

extern "C" int crimson_form_sync(crimson_form *f) {
  if (!f) {
    crimson::set_error("crimson_form_sync: null handle");
    return CRIMSON_EARG;
  }
  CR_HIP(hipStreamSynchronize(form_stream(f)));
  return CRIMSON_OK;
}

extern "C" int crimson_elmpvsq_vjp_dev(crimson_form *f, int nface, const int *d_glob, const double *d_nabi,
                                       double c1, const double *d_y, const double *d_w4, double *d_ybar,
                                       double *d_cbar) {
  if (!f || nface < 0 || (nface > 0 && (!d_glob || !d_nabi)) || !d_y || !d_w4 || !d_ybar || !d_cbar) {
    crimson::set_error("crimson_elmpvsq_vjp_dev: invalid arguments");
    return CRIMSON_EARG;
  }
  const int N = form_nshg(f);
  hipStream_t s = form_stream(f);
  pvsq_vjp_dot_kernel<<<1, 256, 0, s>>>(nface, d_glob, d_nabi, d_y, N, d_w4, d_cbar);
  if (nface > 0)
    pvsq_vjp_axpy_kernel<<<crimson::grid_for(nface, 256), 256, 0, s>>>(nface, d_glob, d_nabi, N, c1, d_cbar,
                                                                         d_ybar);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}
