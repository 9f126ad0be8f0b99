// Shared internals of libcrimson_amd: error reporting, device buffers and the
// memLS matrix handle that both the formation and the solver touch.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/crimson_amd.h"

namespace crimson {

void set_error(const char *fmt, ...);
const std::string &last_error();

#define CR_HIP(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      crimson::set_error("%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                               \
      return CRIMSON_EHIP;                                                        \
    }                                                                             \
  } while (0)

#define CR_TRY(expr)            \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != CRIMSON_OK) return rc_; \
  } while (0)

// Serial numbers of device allocations: every successful allocation gets a
// new one, so a captured HIP graph can tell whether a buffer it reads has
// been released or replaced since the capture (even at the same address).
inline unsigned long long next_alloc_serial() {
  static std::atomic<unsigned long long> counter{0};
  return ++counter;
}

// Owning device allocation.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  unsigned long long serial = 0;  // 0 while nothing is allocated
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), serial(o.serial) {
    o.p = nullptr;
    o.n = 0;
    o.serial = 0;
  }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      serial = o.serial;
      o.p = nullptr;
      o.n = 0;
      o.serial = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    serial = 0;
  }
  int alloc(size_t count) {
    if (count == n && p) return CRIMSON_OK;
    release();
    if (count == 0) return CRIMSON_OK;
    hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
      return CRIMSON_ENOMEM;
    }
    n = count;
    serial = next_alloc_serial();
    return CRIMSON_OK;
  }
  int upload(const T *h, size_t count, hipStream_t s) {
    CR_TRY(alloc(count));
    if (count) CR_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    return CRIMSON_OK;
  }
};

// Workgroup index b of nb remapped so that each XCD (blocks are dealt
// round-robin over the 8 XCDs) walks a contiguous range of rows or elements
// (chunk 0) and keeps the vector entries or nodes it gathers in its own L2;
// or with the workgroups dealt to the XCDs in turns of `chunk` consecutive
// ones (XCD x takes chunks x, x + 8, x + 16, ...), so that the XCDs work on
// neighbouring ranges at the same time and data shared by items far apart
// in the order (a layer of a structured numbering) is found again in the
// shared MALL instead of HBM.
__device__ inline unsigned xcd_block(unsigned chunk, unsigned b, unsigned nb) {
  if (chunk == 0) {
    const unsigned per = nb / 8u;
    if (per == 0 || b >= per * 8u) return b;
    return (b % 8u) * per + b / 8u;
  }
  const unsigned full = nb / (8u * chunk) * (8u * chunk);
  if (b >= full) return b;
  const unsigned x = b % 8u, k = b / 8u;  // the k-th workgroup dealt to XCD x
  return ((k / chunk) * 8u + x) * chunk + k % chunk;
}
__device__ inline unsigned xcd_block(unsigned chunk) { return xcd_block(chunk, blockIdx.x, gridDim.x); }
__device__ inline unsigned xcd_block() { return xcd_block(0u); }

// Number of workgroups of `threads` covering n items.
inline unsigned grid_for(size_t n, unsigned threads) {
  return (unsigned)((n + threads - 1) / threads);
}

}  // namespace crimson

struct crimson_form;
struct crimson_lhs;
namespace crimson {
hipStream_t form_stream(const crimson_form *f);
int form_nshg(const crimson_form *f);
crimson_form_params form_params_of(const crimson_form *f);
bool form_has_wall(const crimson_form *f);  // crimson_form_set_wall holds
// The wall displacement's steps inside a captured timestep (no-ops without
// a wall): predict u from uold, uAlpha, correct u, uold = u at the end.
// A wall needs crimson_form_set_wall_state_dev first (CRIMSON_ESTATE).
int form_wall_step_ready(crimson_form *f);
int form_wall_enqueue_predict(crimson_form *f, const double *d_yold, const double *d_acold);
int form_wall_enqueue_alpha(crimson_form *f);
int form_wall_enqueue_correct(crimson_form *f, const double *d_Res4);
int form_wall_enqueue_update(crimson_form *f);

// What a graph of the solver side adds to a replay: the state it holds on
// the device, read back after each launch, and the buffers and settings it
// baked in (a replay is refused once they change).
struct GraphExtra {
  virtual ~GraphExtra() = default;
  virtual int after_launch() { return CRIMSON_OK; }
  virtual std::vector<unsigned long long> signature() const { return {}; }
};

// The formation's graph signature (form.hip): the allocation serial of every
// buffer a captured formation reads or writes and the settings that chose
// its kernels.
std::vector<unsigned long long> graph_signature(const crimson_form *f, const crimson_lhs *lhs);
// Pre-capture set-up of a formation into lhs's matrix (its allocation and
// the assembly schedule); then form_enqueue_to_lhs enqueues that formation
// on the handle's stream without any synchronisation (capturable), and
// form_enqueue_bc3res bc3Res on Res4.
int form_prepare_to_lhs(crimson_form *f, const double *d_x, crimson_lhs *lhs);
int form_enqueue_to_lhs(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x, double *d_Res4,
                        crimson_lhs *lhs);
int form_enqueue_bc3res(crimson_form *f, double *d_res, int res4_layout);
}  // namespace crimson

// A captured HIP graph (crimson_form_graph_create, crimson_ls_graph_create,
// crimson_step_graph_create).
struct crimson_graph {
  hipGraph_t g = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t s = nullptr;
  const crimson_form *f = nullptr;  // with lhs: the formation's signature applies
  const crimson_lhs *lhs = nullptr;
  std::vector<unsigned long long> sig;  // the signature at capture
  crimson::GraphExtra *extra = nullptr;  // owned
  // a capture cut into segments: graphs launched back to back on s, with no
  // host decision between them (g and exec unused then)
  std::vector<hipGraph_t> seg;
  std::vector<hipGraphExec_t> segx;
};

// memLS face (memLS_faceType, memLS.f90:83-100), single task.
struct CrFace {
  int nNo = 0, dof = 0, bGrp = CRIMSON_BC_DIR;
  bool created = false, coupled = false, inc = true;
  double res = 0.0, nS = 0.0;
  std::vector<int> glob;      // 0-based node
  std::vector<double> val;    // (dof, nNo)
  crimson::DBuf<int> d_glob;
  crimson::DBuf<double> d_val;   // (dof, nNo)
  crimson::DBuf<double> d_valM;  // (dof, nNo) = val * W
  // the columns j of the rows of the face's nodes (sorted) and, for each,
  // its (face node a, CSR entry k) pairs in (a, k) order: the support of
  // the face dot as a functional of the G product's operand (built once
  // per face)
  bool qpat = false;
  int nsup = 0;
  crimson::DBuf<int> d_sup, d_qptr, d_qa, d_qk;
  crimson::DBuf<double> d_q;  // the functional's coefficients, per solve
};

// memLS matrix handle.  The matrix of block size vdof (1..4) is held as
// vdof*vdof planes of nnz doubles, plane m = row-major entry (m / vdof,
// m % vdof) of every block, i.e. Val(m+1, :) of the reference layout.
struct crimson_lhs {
  int nNo = 0, nnz = 0, nFaces = 0;
  hipStream_t stream = nullptr;
  std::vector<int> h_rowptr, h_col, h_diag, h_tr;  // 0-based
  crimson::DBuf<int> d_rowptr, d_col, d_diag, d_tr;
  // the column of each block as its offset from the row (col - row), when
  // every offset fits 16 bits: the products read 2 bytes per block instead
  // of 4 (empty otherwise)
  crimson::DBuf<short> d_col16;
  int vdof = 0;                   // block size of the stored matrix
  int prec = 64;                  // 64: d_V holds the matrix; 32: d_V32 does
  crimson::DBuf<double> d_V;      // planes [vdof*vdof][nnz]
  crimson::DBuf<float> d_V32;     // FP32 planes [vdof*vdof][nnz]
  bool gt_is_d = false;           // Gt == mD planes (formation-produced)
  crimson::DBuf<double> d_Gt;     // [3][nnz] when !gt_is_d
  crimson::DBuf<float> d_Gt32;    // the same for an FP32 matrix
  std::vector<CrFace> face;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
