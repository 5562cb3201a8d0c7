// mhx_internal.hpp — shared declarations of the mhx engine (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mhx.h"
#include "expr.hpp"

namespace mhx {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};

#define MHX_HIP(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::mhx::Error{MHX_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define MHX_REQUIRE(cond, code, text)                     \
  do {                                                    \
    if (!(cond)) throw ::mhx::Error{(code), (text)};      \
  } while (0)

// ---------------------------------------------------------------------------
// 1-D tables passed by value in kernel arguments (kernarg segment, read with
// scalar loads when the index is wave-uniform).
// ---------------------------------------------------------------------------
constexpr int kMaxP1 = 8;   // p + 1
constexpr int kMaxQ1 = 8;   // Gauss points per axis

struct Tables1D {
  double B[kMaxP1][kMaxQ1];  // Lagrange value  L_i(xi_q), equispaced nodes
  double D[kMaxP1][kMaxQ1];  // Lagrange slope  L_i'(xi_q)
  double N[2][kMaxQ1];       // linear geometry map values
  double M[2][kMaxQ1];       // linear geometry map slopes
  double W[kMaxQ1];          // Gauss weights
};

// Structured box mesh description used by the kernels (all arithmetic,
// no connectivity tables: element (ex,ey,ez) owns lattice nodes
// p*e + {0..p} per axis; rows are lexicographic in the node lattice).
struct BoxMesh {
  int dim, p, nq1;
  int nf;         // unknowns per node; DOF = node * nf + field
  int nel[3];     // elements per axis
  int nn[3];      // DOF nodes per axis (p*nel+1)
  int nvx[3];     // vertices per axis (nel+1)
  long ne, ndof, nvert, nnode;
};

// Thermal coefficients (compiled form of "thermal diffusion" k0_e*(1+a*T^2)
// and "thermal source" src_e + sum_m amp_m prod_d s(w_md x_d)).
struct ThermalCoef {
  const double* k0;   // [ne] device, or nullptr (== 1)
  const double* src;  // [ne] device, or nullptr (== 0)
  double diff_u2;
  int n_modes;
  double amp[MHX_MAX_MODES];
  double freq[MHX_MAX_MODES][3];
};

struct AsmArgs {
  BoxMesh m;
  ThermalCoef coef;
  const double* vx;       // [nvert*3] vertex coordinates
  const double* u;        // [ndof]
  const int32_t* rowptr;  // [ndof+1]
  const uint8_t* fixed;   // [ndof]
  double* res;            // [ndof] or nullptr
  double* jac;            // [nnz]  or nullptr
  int color;              // colour bits (cx | cy<<1 | cz<<2): element or block colour
  long ncolor;            // elements (element colouring) / blocks (block colouring) of this colour
  int cn[3];              // per-axis count inside this colour
  int blocked;            // 0: element colouring, 1: block colouring, 2: x-segments
  int nseg;               // segments per x-row (segment scheduling)
  int nsub;               // sub-phases per workgroup (1, or 2^dim element colours)
  int gpb;                // element groups per workgroup
  int s[3];               // block colouring: groups per axis (prod = gpb)
  int bw[3];              // block colouring: block width in elements (2*s)
  int first_touch;        // overwrite on first touch instead of accumulate
  // transient (backward differences): dT/dt = alpha_t u + sum_j hist_j phi_j
  const double* hist;     // [ndof] nodal beta_t = sum_k BDF_k u_prev[k-1] / (dt b_s), or nullptr (steady)
  double alpha_t;         // BDF_0 / (dt b_s)
  const double* hist_u;   // [ndof] nodal beta_u of a Butcher stage, or nullptr (alpha_u = 1, beta_u = 0)
  double alpha_u;         // A_ss / b_s
  double rho_cp;          // density * specific heat
  ExprSet xp;             // compiled coefficient functions (xp.active = 0: the closed forms)
};

// compiled coefficient functions of a context, host side (expr.hip)
struct ExprHost {
  std::vector<ExprIns> code;
  std::vector<double> consts;
  int off[XF_COUNT] = {};
  int len[XF_COUNT] = {};
  double dflt[XF_COUNT] = {};
  int active = 0;
};
void expr_compile(const mhx_problem_desc& d, ExprHost& out);

enum ScatterMode { kElementColouring = 0, kBlockColouring = 1, kSegmentMerge = 2, kPencilStream = 3 };

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count) {
      hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
      if (e != hipSuccess) {
        p = nullptr;
        throw Error{MHX_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e)};
      }
    }
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// ---------------------------------------------------------------------------
// phase timers (named as the reference's Teuchos timers)
// ---------------------------------------------------------------------------
struct Profiler {
  bool on = false;
  struct Phase {
    std::string name;
    double ms = 0.0;
    long count = 0;
  };
  std::vector<Phase> phases;
  struct Pending {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  int phase_id(const char* name);
  hipEvent_t get();
  void begin(const char* name, hipStream_t s, Pending& out);
  void end(Pending& p, hipStream_t s);
  void drain();
  ~Profiler();
};

struct ScopedPhase {
  Profiler& P;
  hipStream_t s;
  Profiler::Pending pend{-1, nullptr, nullptr};
  ScopedPhase(Profiler& p, const char* name, hipStream_t st) : P(p), s(st) {
    if (P.on) P.begin(name, s, pend);
  }
  ~ScopedPhase() {
    if (P.on && pend.phase >= 0) P.end(pend, s);
  }
};

// ---------------------------------------------------------------------------
// Krylov workspace
// ---------------------------------------------------------------------------
template <class T>
struct KrylovVecs {
  DevBuf<T> r, z, p, q, tmp, tmp2, V, prec;  // prec: 1/diag or nodal block inverses
  DevBuf<T> rhs, sol;                         // inner system of the refinement loop
};
struct KrylovWork {
  KrylovVecs<double> d;
  KrylovVecs<float> f;
  DevBuf<float> jac32;
  DevBuf<double> x64;  // FP64 residual / correction of the refinement loop
  DevBuf<double> part, scal;
  DevBuf<double> cgs;  // CG scalars on the device (rho, p.q, ||r||^2 window)
};

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct Amg;  // smoothed-aggregation hierarchy (amg.hip)

struct Ctx {
  mhx_problem_desc desc{};
  BoxMesh mesh{};
  Tables1D tab{};
  int nb = 0, nq = 0;
  long nnz = 0;
  std::vector<int32_t> h_rowptr, h_colidx;
  std::vector<uint8_t> h_fixed;
  std::vector<double> h_dbc;   // [ndof] Dirichlet values (0 elsewhere)
  std::vector<double> h_vx;    // [nvert*3]
  std::vector<int32_t> h_dbc_list;
  // device
  DevBuf<int32_t> rowptr, colidx, dbc_list;
  DevBuf<uint8_t> fixed;
  DevBuf<double> vx, k0, src, dbc_vals;
  DevBuf<double> u, res, jac, du, aux;      // context-owned device state (mhx_state_ptr)
  DevBuf<double> stage_u, stage_r, stage_j;  // staging for host-pointer calls
  DevBuf<double> stage_a, stage_b;
  DevBuf<double> norm_buf;
  bool jac_valid = false;
  KrylovWork kw;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  Profiler prof;
  int spmv_lanes = 16;
  int lattice_cols = 1;  // sparse products take columns from the lattice closed form (0: from colidx)
  int amg_smoother = -1;  // -1 automatic, 0 Chebyshev(2) on the point diagonal, 1 nodal block Jacobi, 2 block Gauss-Seidel
  int amg_reuse = 0;       // 1: later AMG set-ups keep the coarse levels and refresh the fine level
  int amg_sweeps = 3;      // block-Jacobi smoother: sweeps before and after the coarse correction
  int amg_omega_pct = 40;  // block-Jacobi smoother: damping in percent
  int nfield = 1;  // unknowns per node (consecutive DOFs): block size of block Jacobi
  int scatter_mode = kElementColouring;
  bool box_mesh = false;  // every element an axis-aligned box (structured box, no node perturbation)
  int q2_segment = 12;    // pencil streaming: elements per x-segment (even)
  std::shared_ptr<Amg> amg;  // built by a solve with the AMG preconditioner, kept for reuse of its buffers
  // time-integration state (mhx_set_time_state); bdf_order 0 = steady
  int bdf_order = 0;
  double alpha_t = 0.0;
  DevBuf<double> hist;       // [ndof] nodal history term of dT/dt
  DevBuf<double> hist_u;     // [ndof] nodal beta_u of a Butcher stage (released for one-stage tableaus)
  double alpha_u = 1.0;
  // compiled coefficient functions
  ExprHost xh;
  DevBuf<ExprIns> x_code;
  DevBuf<double> x_consts;
  std::string err;
};

// mesh / graph (mesh.hip)
void build_box(Ctx& c);
void gauss_legendre_desc(int n, double* x, double* w);
void lagrange_equispaced(int p, double xi, double* L, double* dL);

// assembly (assemble.hip)
bool thermal_supported(int dim, int p, int nq1);
// Q2 hexahedra on box meshes: pencil-streamed pair-factorised assembly (assemble_q2box.hip)
bool q2box_applicable(const Ctx& c);
void launch_q2box(Ctx& c, AsmArgs& a, bool jac);
// thermal assembly with compiled coefficient functions, element colouring
// (assemble_expr.hip)
void launch_thermal_expr(Ctx& c, AsmArgs& a, bool jac);
void assemble_thermal(Ctx& c, const double* u, double* res, double* jac, bool want_res, bool want_jac,
                      bool accumulate);
// thermal Neumann boundary residual (bc.hip)
void neumann_residual(Ctx& c, double* res);
// thermal weak (Nitsche) Dirichlet sides: residual and Jacobian terms added
// after the volume assembly (bc.hip)
bool nitsche_supported(int dim, int p);
void nitsche_boundary(Ctx& c, const double* u, double* res, double* jac);
// navier stokes assembly (ns.hip)
bool ns_supported(int dim, int p, int nq1);
void assemble_ns(Ctx& c, const double* u, double* res, double* jac, bool want_res, bool want_jac, bool accumulate);

// linear algebra (krylov.hip)
void spmv(Ctx& c, const double* A, const double* x, double* y);
double dot(Ctx& c, const double* a, const double* b, long n);
// a . b reduced into *slot on the device (no host copy, no synchronisation)
void dot_dev(Ctx& c, const double* a, const double* b, long n, double* slot);
double norm2(Ctx& c, const double* a, long n);
double norm_inf(Ctx& c, const double* a, long n);
int krylov_solve(Ctx& c, const double* A, const double* b, double* x, const mhx_krylov_opts& o,
                 mhx_solve_stats* st);
// CG / Jacobi-PCG with every device buffer in a static per-GPU array (no allocation)
int krylov_solve_noalloc(Ctx& c, const double* A, const double* b, double* x, const mhx_krylov_opts& o,
                         mhx_solve_stats* st);
void set_dirichlet_dev(Ctx& c, double* u);
void axpy(Ctx& c, long n, double a, const double* x, double* y);

// MatrixMarket files (mmio.hip); dev: the pointer is a device pointer
void write_matrix_market(Ctx& c, const char* path, const double* vals, bool dev);
void write_vector_market(Ctx& c, const char* path, const double* vec, bool dev);
void read_matrix_market(Ctx& c, const char* path, double* vals, bool dev);
void read_vector_market(Ctx& c, const char* path, double* vec, bool dev);

// smoothed-aggregation AMG (amg.hip): set-up from the FP64 operator A on the
// context's pattern (and FP32 copies when fp32), and one V-cycle out = M^-1 v
// with level-0 operator A (the FP64 A, or its FP32 copy)
void amg_setup(Ctx& c, const double* A, bool fp32);
void amg_apply(Ctx& c, const double* A, const double* v, double* out);
void amg_apply(Ctx& c, const float* A, const float* v, float* out);
int amg_levels(const Ctx& c);
double amg_lambda(const Ctx& c, int level);
void amg_download(Ctx& c, int level, int which, int64_t* rows, int64_t* cols, int64_t* nnz, int32_t* rp, int32_t* ci,
                  double* vals);

}  // namespace mhx
