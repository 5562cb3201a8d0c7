// capi.hip — the C ABI (include/mhx.h) over the engine: context setup,
// assembly, linear and Newton solves, error norms, profiling, seeded inputs.
// Nothing here throws across the boundary; failures become negative status
// codes with a message in mhx_last_error().

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "mhx_internal.hpp"

namespace mhx {

static thread_local std::string g_last_error;

// ---- profiler ----------------------------------------------------------------
int Profiler::phase_id(const char* name) {
  for (size_t i = 0; i < phases.size(); ++i)
    if (phases[i].name == name) return (int)i;
  phases.push_back(Phase{name, 0.0, 0});
  return (int)phases.size() - 1;
}
hipEvent_t Profiler::get() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e;
  MHX_HIP(hipEventCreate(&e));
  return e;
}
void Profiler::begin(const char* name, hipStream_t s, Pending& out) {
  out.phase = phase_id(name);
  out.a = get();
  out.b = get();
  MHX_HIP(hipEventRecord(out.a, s));
}
void Profiler::end(Pending& p, hipStream_t s) {
  (void)hipEventRecord(p.b, s);
  pending.push_back(p);
  if (pending.size() > 4096) drain();
}
void Profiler::drain() {
  for (Pending& p : pending) {
    (void)hipEventSynchronize(p.b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    phases[p.phase].ms += ms;
    phases[p.phase].count += 1;
    pool.push_back(p.a);
    pool.push_back(p.b);
  }
  pending.clear();
}
Profiler::~Profiler() {
  drain();
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

// ---- helpers -----------------------------------------------------------------
// compiled coefficient functions to the device (code and constants)
static void upload_functions(Ctx& c) {
  if (!c.xh.active) {
    c.x_code.release();
    c.x_consts.release();
    return;
  }
  c.x_code.alloc(c.xh.code.size());
  MHX_HIP(hipMemcpyAsync(c.x_code.p, c.xh.code.data(), c.xh.code.size() * sizeof(ExprIns), hipMemcpyHostToDevice,
                         c.stream));
  c.x_consts.alloc(c.xh.consts.empty() ? 1 : c.xh.consts.size());
  if (!c.xh.consts.empty())
    MHX_HIP(hipMemcpyAsync(c.x_consts.p, c.xh.consts.data(), c.xh.consts.size() * sizeof(double),
                           hipMemcpyHostToDevice, c.stream));
  MHX_HIP(hipStreamSynchronize(c.stream));
}

static void validate(const mhx_problem_desc* d) {
  MHX_REQUIRE(d, MHX_ERR_ARG, "null problem description");
  MHX_REQUIRE(d->n_functions >= 0 && d->n_functions <= 256, MHX_ERR_ARG, "n_functions out of range");
  MHX_REQUIRE(d->dim == 2 || d->dim == 3, MHX_ERR_ARG, "dim must be 2 or 3");
  MHX_REQUIRE(d->order >= 1 && d->order <= kMaxP1 - 1, MHX_ERR_ARG, "order out of range");
  MHX_REQUIRE(d->quadrature >= 0 && d->quadrature / 2 + 1 <= kMaxQ1, MHX_ERR_ARG, "quadrature out of range");
  MHX_REQUIRE(d->nx >= 1 && d->ny >= 1 && (d->dim == 2 || d->nz >= 1), MHX_ERR_ARG, "element counts must be >= 1");
  MHX_REQUIRE(d->xmax > d->xmin && d->ymax > d->ymin && (d->dim == 2 || d->zmax > d->zmin), MHX_ERR_ARG,
              "empty box");
  MHX_REQUIRE(d->n_modes >= 0 && d->n_modes <= MHX_MAX_MODES, MHX_ERR_ARG, "n_modes out of range");
  MHX_REQUIRE(d->n_nq_modes >= 0 && d->n_nq_modes <= MHX_MAX_NQ_MODES, MHX_ERR_ARG, "n_nq_modes out of range");
  MHX_REQUIRE(d->th_density >= 0.0 && d->th_specific_heat >= 0.0, MHX_ERR_ARG,
              "density and specific heat must be >= 0 (0 = 1.0)");
  if (d->wdbc_sides) {
    MHX_REQUIRE(d->physics == MHX_PHYS_THERMAL, MHX_ERR_UNSUPPORTED,
                "weak Dirichlet sides are built for the thermal module only");
    MHX_REQUIRE(nitsche_supported(d->dim, d->order), MHX_ERR_UNSUPPORTED,
                "weak Dirichlet sides: no kernel built for this dimension and order");
    MHX_REQUIRE((d->wdbc_sides & (d->dbc_sides | d->neumann_sides)) == 0, MHX_ERR_ARG,
                "a side cannot be weak Dirichlet and strong Dirichlet or Neumann at once");
    MHX_REQUIRE(d->wdbc_form >= 0 && d->wdbc_form <= 2, MHX_ERR_ARG, "wdbc_form must be 0, 1 or 2");
    MHX_REQUIRE(d->n_wd_modes >= 0 && d->n_wd_modes <= MHX_MAX_NQ_MODES, MHX_ERR_ARG, "n_wd_modes out of range");
  }
  MHX_REQUIRE(d->neumann_sides == 0 || d->physics == MHX_PHYS_THERMAL || d->physics == MHX_PHYS_NAVIER_STOKES,
              MHX_ERR_UNSUPPORTED, "Neumann sides are built for the thermal and Navier-Stokes modules only");
  if (d->physics == MHX_PHYS_NAVIER_STOKES) {
    MHX_REQUIRE(d->ns_rho > 0.0 && d->ns_mu > 0.0, MHX_ERR_ARG, "navier stokes: density and viscosity must be > 0");
    MHX_REQUIRE(ns_supported(d->dim, d->order, d->quadrature / 2 + 1), MHX_ERR_UNSUPPORTED,
                "navier stokes: Q1-Q1 with quadrature 2 or 3 only");
    return;
  }
  MHX_REQUIRE(d->physics == MHX_PHYS_THERMAL, MHX_ERR_UNSUPPORTED, "physics module not built in this library");
  MHX_REQUIRE(thermal_supported(d->dim, d->order, d->quadrature / 2 + 1), MHX_ERR_UNSUPPORTED,
              "no assembly kernel for this dim/order/quadrature");
}

// physics dispatch (the reference's per-module volumeResidual, physicsImporter.cpp:76-272)
static void assemble_physics(Ctx& c, const double* u, double* res, double* jac, bool want_res, bool want_jac,
                             bool accumulate) {
  if (c.desc.physics == MHX_PHYS_NAVIER_STOKES) assemble_ns(c, u, res, jac, want_res, want_jac, accumulate);
  else assemble_thermal(c, u, res, jac, want_res, want_jac, accumulate);
}

// copy-in / copy-out of caller vectors for host-pointer calls
struct VecArg {
  Ctx& c;
  const double* user;
  double* dev;
  bool device;
  size_t n;
  VecArg(Ctx& cc, const double* p, size_t count, bool is_dev, DevBuf<double>& stage, bool copy_in)
      : c(cc), user(p), dev(nullptr), device(is_dev), n(count) {
    if (device) {
      dev = const_cast<double*>(p);
    } else {
      stage.alloc(std::max<size_t>(n, 1));
      dev = stage.p;
      if (copy_in && p) MHX_HIP(hipMemcpyAsync(dev, p, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    }
  }
  void copy_out(double* out) {
    if (!device && out) MHX_HIP(hipMemcpyAsync(out, dev, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  }
};

// ---- error norms at the cubature points (computeError L2 / GRAD) --------------
struct ErrArgs {
  BoxMesh m;
  Tables1D t;
  const double* vx;
  const double* u;
  int nm;
  double amp[MHX_MAX_MODES];
  double freq[MHX_MAX_MODES][3];
  double* part;  // [2 * gridDim]
};

__global__ __launch_bounds__(256) void k_l2err(ErrArgs a) {
  const int dim = a.m.dim, p = a.m.p, p1 = p + 1, nq1 = a.m.nq1;
  const int nb = dim == 3 ? p1 * p1 * p1 : p1 * p1;
  const int nq = dim == 3 ? nq1 * nq1 * nq1 : nq1 * nq1;
  const int nv = 1 << dim;
  double e0 = 0.0, e1 = 0.0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < a.m.ne; e += (long)gridDim.x * blockDim.x) {
    int ec[3] = {(int)(e % a.m.nel[0]), (int)((e / a.m.nel[0]) % a.m.nel[1]), (int)(e / ((long)a.m.nel[0] * a.m.nel[1]))};
    double X[8][3];
    for (int v = 0; v < nv; ++v) {
      long vid = (long)(ec[0] + (v & 1)) + (long)a.m.nvx[0] * ((long)(ec[1] + ((v >> 1) & 1)) + (long)a.m.nvx[1] * (long)(ec[2] + ((v >> 2) & 1)));
      for (int d = 0; d < 3; ++d) X[v][d] = a.vx[vid * 3 + d];
    }
    for (int q = 0; q < nq; ++q) {
      int qi[3] = {q % nq1, (q / nq1) % nq1, dim == 3 ? q / (nq1 * nq1) : 0};
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, x[3] = {0, 0, 0};
      for (int v = 0; v < nv; ++v) {
        int vl[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
        double nvv = 1.0;
        for (int d = 0; d < dim; ++d) nvv *= a.t.N[vl[d]][qi[d]];
        for (int r = 0; r < dim; ++r) {
          x[r] += nvv * X[v][r];
          for (int s = 0; s < dim; ++s) {
            double dn = 1.0;
            for (int d = 0; d < dim; ++d) dn *= (d == s) ? a.t.M[vl[d]][qi[d]] : a.t.N[vl[d]][qi[d]];
            J[r][s] += X[v][r] * dn;
          }
        }
      }
      double Ji[3][3], det;
      if (dim == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        Ji[0][0] = J[1][1] / det; Ji[0][1] = -J[0][1] / det; Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
      } else {
        double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        Ji[0][0] = c00 / det; Ji[1][0] = c01 / det; Ji[2][0] = c02 / det;
        Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
      }
      double w = fabs(det);
      for (int d = 0; d < dim; ++d) w *= a.t.W[qi[d]];
      double Th = 0.0, gr[3] = {0, 0, 0};
      for (int j = 0; j < nb; ++j) {
        int jl[3] = {j % p1, (j / p1) % p1, dim == 3 ? j / (p1 * p1) : 0};
        long node = (long)(p * ec[0] + jl[0]) + (long)a.m.nn[0] * ((long)(p * ec[1] + jl[1]) + (long)a.m.nn[1] * (long)(p * ec[2] + jl[2]));
        double uj = a.u[node];
        double b = 1.0;
        for (int d = 0; d < dim; ++d) b *= a.t.B[jl[d]][qi[d]];
        Th += uj * b;
        for (int c = 0; c < dim; ++c) {
          double gg = 1.0;
          for (int d = 0; d < dim; ++d) gg *= (d == c) ? a.t.D[jl[d]][qi[d]] : a.t.B[jl[d]][qi[d]];
          gr[c] += uj * gg;
        }
      }
      double gp[3] = {0, 0, 0};
      for (int c = 0; c < dim; ++c)
        for (int b = 0; b < dim; ++b) gp[c] += Ji[b][c] * gr[b];
      double Tt = 0.0, gt[3] = {0, 0, 0};
      for (int mm = 0; mm < a.nm; ++mm) {
        double s[3], cc[3];
        for (int d = 0; d < dim; ++d) {
          double fw = a.freq[mm][d];
          s[d] = fw != 0.0 ? sin(fw * x[d]) : 1.0;
          cc[d] = fw != 0.0 ? fw * cos(fw * x[d]) : 0.0;
        }
        double t = a.amp[mm];
        for (int d = 0; d < dim; ++d) t *= s[d];
        Tt += t;
        for (int c = 0; c < dim; ++c) {
          double g = a.amp[mm];
          for (int d = 0; d < dim; ++d) g *= (d == c) ? cc[d] : s[d];
          gt[c] += g;
        }
      }
      e0 += (Th - Tt) * (Th - Tt) * w;
      for (int c = 0; c < dim; ++c) e1 += (gp[c] - gt[c]) * (gp[c] - gt[c]) * w;
    }
  }
  __shared__ double sh[2][256];
  sh[0][threadIdx.x] = e0;
  sh[1][threadIdx.x] = e1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.part[2 * blockIdx.x] = sh[0][0];
    a.part[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// plain copy, two doubles per thread per iteration (calibration kernel)
__global__ __launch_bounds__(256) void k_copy_probe(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t n2 = n / 2;
  const double2* x2 = reinterpret_cast<const double2*>(x);
  double2* y2 = reinterpret_cast<double2*>(y);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    y2[i] = x2[i];
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) y[n - 1] = x[n - 1];
}

// ---- Newton (SolverManager::nonlinearSolver, solverManager_solvers.hpp:290-643) --
static int newton(Ctx& c, double* u, const mhx_newton_opts& o, mhx_solve_stats* st) {
  const long n = c.mesh.ndof;
  c.res.alloc(n);
  c.jac.alloc(c.nnz);
  c.aux.alloc(n);
  c.du.alloc(n);
  DevBuf<double>& du = c.du;
  MHX_HIP(hipMemsetAsync(du.p, 0, n * sizeof(double), c.stream));
  double first = 0.0, alpha = 1.0;
  int it = 0, status = MHX_OK;
  bool proceed = true;
  const int maxit = o.max_iters > 0 ? o.max_iters : 10;
  const double tol = o.nl_tol > 0 ? o.nl_tol : 1e-6;
  const double abstol = o.abs_tol > 0 ? o.abs_tol : std::min(tol, 1.0e-6);
  while (proceed) {
    assemble_physics(c, u, c.res.p, nullptr, true, false, false);  // assembleRes (ScalarT)
    const double nrm = norm_inf(c, c.res.p, n);
    if (!std::isfinite(nrm)) throw Error{MHX_ERR_NONFINITE, "Newton: non-finite residual norm"};
    double scaled;
    if (it == 0) {
      first = nrm;
      scaled = 1.0;
    } else {
      scaled = nrm / first;
    }
    if (st) {
      if (it == 0) st->res_norm0 = nrm;
      st->res_norm = nrm;
    }
    bool solve = true;
    if (o.backtracking && scaled > 1.1) {
      solve = false;
      alpha *= 0.5;
      axpy(c, n, -alpha, du.p, u);
    } else if (!o.relative_off) {
      if (scaled < tol || nrm < 1.0e-100) {
        solve = false;
        proceed = false;
      }
    } else if (o.use_abs_tol && nrm < abstol) {
      solve = false;
      proceed = false;
    }
    if (solve) {
      assemble_physics(c, u, c.aux.p, c.jac.p, true, true, false);  // assembleJacRes (AD)
      c.jac_valid = true;
      MHX_HIP(hipMemsetAsync(du.p, 0, n * sizeof(double), c.stream));
      int s = krylov_solve(c, c.jac.p, c.res.p, du.p, o.linear, st);
      if (s != MHX_OK) status = s;
      alpha = 1.0;
      axpy(c, n, alpha, du.p, u);
    }
    ++it;
    if (it >= maxit) proceed = false;
  }
  if (st) st->nl_iters = it;
  MHX_HIP(hipStreamSynchronize(c.stream));
  return status;
}

}  // namespace mhx

using namespace mhx;

#define MHX_GUARD_BEGIN try {
#define MHX_GUARD_END(ctxp)                          \
  }                                                  \
  catch (const Error& e) {                           \
    g_last_error = e.msg;                            \
    if (ctxp) (ctxp)->err = e.msg;                   \
    return e.code;                                   \
  }                                                  \
  catch (const std::exception& e) {                  \
    g_last_error = e.what();                         \
    if (ctxp) (ctxp)->err = e.what();                \
    return MHX_ERR_ARG;                              \
  }

struct mhx_ctx : public Ctx {};

extern "C" {

const char* mhx_version(void) { return "mhx 0.1 (gfx950, fp64 assembly, sum-factorised AD columns)"; }

const char* mhx_last_error(const mhx_ctx* ctx) {
  if (ctx && !ctx->err.empty()) return ctx->err.c_str();
  return g_last_error.c_str();
}

int mhx_ctx_create(const mhx_problem_desc* desc, mhx_ctx** out) {
  mhx_ctx* c = nullptr;
  MHX_GUARD_BEGIN
  MHX_REQUIRE(out, MHX_ERR_ARG, "null output pointer");
  *out = nullptr;
  validate(desc);
  int ndev = 0;
  MHX_HIP(hipGetDeviceCount(&ndev));
  MHX_REQUIRE(ndev > 0, MHX_ERR_DEVICE, "no HIP device");
  c = new mhx_ctx();
  c->desc = *desc;
  // coefficient functions: compiled before any device work
  expr_compile(*desc, c->xh);
  MHX_REQUIRE(!c->xh.active || desc->physics == MHX_PHYS_THERMAL, MHX_ERR_UNSUPPORTED,
              "coefficient functions are built for the thermal module");
  MHX_REQUIRE(!c->xh.active || !desc->wdbc_sides, MHX_ERR_UNSUPPORTED,
              "weak Dirichlet sides take the closed-form diffusion; not built with coefficient functions");
  // arrays in the description are consumed at create time only
  MHX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->own_stream = true;
  build_box(*c);
  upload_functions(*c);
  // the lattice without node perturbation has axis-aligned box elements: Q2
  // thermal assembly then defaults to the pencil-streamed kernel
  c->box_mesh = desc->perturb == 0.0;
  if (q2box_applicable(*c)) c->scatter_mode = kPencilStream;
  // Navier-Stokes: x-segments merged in LDS (each row's x-runs written once
  // per segment) by default
  if (desc->physics == MHX_PHYS_NAVIER_STOKES) c->scatter_mode = kSegmentMerge;
  c->desc.k0 = nullptr;
  c->desc.src_elem = nullptr;
  c->desc.dbc_values = nullptr;
  c->desc.function_names = nullptr;
  c->desc.function_exprs = nullptr;
  *out = c;
  return MHX_OK;
  }
  catch (const Error& e) {
    g_last_error = e.msg;
    if (c) {
      if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
      delete c;
    }
    return e.code;
  }
}

void mhx_ctx_destroy(mhx_ctx* ctx) {
  if (!ctx) return;
  (void)hipStreamSynchronize(ctx->stream);
  ctx->prof.drain();
  hipStream_t s = ctx->stream;
  bool own = ctx->own_stream;
  delete ctx;
  if (s && own) (void)hipStreamDestroy(s);
}

int mhx_sizes(const mhx_ctx* ctx, int64_t* n_dof, int64_t* nnz, int64_t* n_elem, int32_t* n_colors) {
  if (!ctx) return MHX_ERR_ARG;
  if (n_dof) *n_dof = ctx->mesh.ndof;
  if (nnz) *nnz = ctx->nnz;
  if (n_elem) *n_elem = ctx->mesh.ne;
  if (n_colors) *n_colors = 1 << ctx->mesh.dim;
  return MHX_OK;
}

int mhx_crs_pattern(const mhx_ctx* cctx, int32_t* rowptr, int32_t* colidx) {
  mhx_ctx* ctx = const_cast<mhx_ctx*>(cctx);
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx, MHX_ERR_ARG, "null ctx");
  if (rowptr) std::copy(ctx->h_rowptr.begin(), ctx->h_rowptr.end(), rowptr);
  if (colidx) {
    MHX_HIP(hipMemcpyAsync(colidx, ctx->colidx.p, ctx->nnz * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    MHX_HIP(hipStreamSynchronize(ctx->stream));
  }
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_dirichlet_mask(const mhx_ctx* ctx, int32_t* mask) {
  if (!ctx || !mask) return MHX_ERR_ARG;
  for (long g = 0; g < ctx->mesh.ndof; ++g) mask[g] = ctx->h_fixed[g];
  return MHX_OK;
}

int mhx_set_dirichlet(mhx_ctx* ctx, double* u, int32_t flags) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && u, MHX_ERR_ARG, "null argument");
  if (flags & MHX_DEVICE_PTRS) {
    set_dirichlet_dev(*ctx, u);
    MHX_HIP(hipStreamSynchronize(ctx->stream));
  } else {
    for (int32_t g : ctx->h_dbc_list) u[g] = ctx->h_dbc[g];
  }
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_dof_coords(const mhx_ctx* ctx, double* xyz) {
  if (!ctx || !xyz) return MHX_ERR_ARG;
  const BoxMesh& m = ctx->mesh;
  for (long g = 0; g < m.ndof; ++g) {
    const long node = g / m.nf;
    long c[3] = {node % m.nn[0], (node / m.nn[0]) % m.nn[1], node / ((long)m.nn[0] * m.nn[1])};
    long ec[3] = {0, 0, 0};
    double xi[3] = {0, 0, 0};
    for (int a = 0; a < m.dim; ++a) {
      ec[a] = std::min<long>(c[a] / m.p, m.nel[a] - 1);
      xi[a] = -1.0 + 2.0 * (double)(c[a] - m.p * ec[a]) / m.p;
    }
    double X[3] = {0, 0, 0};
    for (int v = 0; v < (1 << m.dim); ++v) {
      long vid = (ec[0] + (v & 1)) + m.nvx[0] * ((ec[1] + ((v >> 1) & 1)) + (long)m.nvx[1] * (ec[2] + ((v >> 2) & 1)));
      double b = 1.0;
      for (int a = 0; a < m.dim; ++a) b *= ((v >> a) & 1) ? 0.5 * (1 + xi[a]) : 0.5 * (1 - xi[a]);
      for (int a = 0; a < 3; ++a) X[a] += b * ctx->h_vx[vid * 3 + a];
    }
    for (int a = 0; a < 3; ++a) xyz[g * 3 + a] = X[a];
  }
  return MHX_OK;
}

int mhx_assemble(mhx_ctx* ctx, const double* u, int32_t flags, double* res, double* jac_vals) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && u, MHX_ERR_ARG, "null argument");
  const bool want_res = (flags & MHX_RES) != 0, want_jac = (flags & MHX_JAC) != 0;
  MHX_REQUIRE(want_res || want_jac, MHX_ERR_ARG, "flags must request MHX_RES and/or MHX_JAC");
  MHX_REQUIRE(!want_res || res, MHX_ERR_ARG, "null residual");
  MHX_REQUIRE(!want_jac || jac_vals, MHX_ERR_ARG, "null jacobian values");
  const bool dev = (flags & MHX_DEVICE_PTRS) != 0;
  const bool acc = (flags & MHX_ACCUMULATE) != 0;
  const long n = ctx->mesh.ndof;
  VecArg U(*ctx, u, n, dev, ctx->stage_u, true);
  VecArg Rv(*ctx, res, want_res ? n : 0, dev, ctx->stage_r, acc);
  VecArg Jv(*ctx, jac_vals, want_jac ? ctx->nnz : 0, dev, ctx->stage_j, acc);
  assemble_physics(*ctx, U.dev, want_res ? Rv.dev : nullptr, want_jac ? Jv.dev : nullptr, want_res, want_jac, acc);
  if (want_res) Rv.copy_out(res);
  if (want_jac) Jv.copy_out(jac_vals);
  if (!dev) MHX_HIP(hipStreamSynchronize(ctx->stream));
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_spmv(mhx_ctx* ctx, const double* jac_vals, const double* x, double* y, int32_t flags) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && x && y, MHX_ERR_ARG, "null argument");
  const bool dev = (flags & MHX_DEVICE_PTRS) != 0;
  const long n = ctx->mesh.ndof;
  const double* A;
  DevBuf<double> jstage;
  if (jac_vals) {
    if (dev) {
      A = jac_vals;
    } else {
      jstage.alloc(ctx->nnz);
      MHX_HIP(hipMemcpyAsync(jstage.p, jac_vals, ctx->nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      A = jstage.p;
    }
  } else {
    MHX_REQUIRE(ctx->jac_valid && ctx->jac.p, MHX_ERR_STATE, "no Jacobian held by the context");
    A = ctx->jac.p;
  }
  VecArg X(*ctx, x, n, dev, ctx->stage_a, true);
  VecArg Y(*ctx, y, n, dev, ctx->stage_b, false);
  spmv(*ctx, A, X.dev, Y.dev);
  Y.copy_out(y);
  MHX_HIP(hipStreamSynchronize(ctx->stream));
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_linear_solve(mhx_ctx* ctx, const double* jac_vals, const double* r, double* x, const mhx_krylov_opts* opts,
                     mhx_solve_stats* stats) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && r && x && opts, MHX_ERR_ARG, "null argument");
  const bool dev = (opts->flags & MHX_DEVICE_PTRS) != 0;
  const long n = ctx->mesh.ndof;
  const double* A;
  DevBuf<double> jstage;
  if (jac_vals) {
    if (dev) {
      A = jac_vals;
    } else {
      jstage.alloc(ctx->nnz);
      MHX_HIP(hipMemcpyAsync(jstage.p, jac_vals, ctx->nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      A = jstage.p;
    }
  } else {
    MHX_REQUIRE(ctx->jac_valid && ctx->jac.p, MHX_ERR_STATE, "no Jacobian held by the context");
    A = ctx->jac.p;
  }
  VecArg R(*ctx, r, n, dev, ctx->stage_a, true);
  VecArg X(*ctx, x, n, dev, ctx->stage_b, true);
  if (stats) std::memset(stats, 0, sizeof(*stats));
  int s = krylov_solve(*ctx, A, R.dev, X.dev, *opts, stats);
  X.copy_out(x);
  MHX_HIP(hipStreamSynchronize(ctx->stream));
  if (stats) stats->status = s;
  return s;
  MHX_GUARD_END(ctx)
}

int mhx_linear_solve_noalloc(mhx_ctx* ctx, const double* jac_vals, const double* r, double* x,
                             const mhx_krylov_opts* opts, mhx_solve_stats* stats) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && r && x && opts, MHX_ERR_ARG, "null argument");
  MHX_REQUIRE((opts->flags & MHX_DEVICE_PTRS) != 0, MHX_ERR_ARG,
              "allocation-free solve: r, x and jac_vals must be device pointers (MHX_DEVICE_PTRS)");
  const double* A = jac_vals;
  if (!A) {
    MHX_REQUIRE(ctx->jac_valid && ctx->jac.p, MHX_ERR_STATE, "no Jacobian held by the context");
    A = ctx->jac.p;
  }
  if (stats) std::memset(stats, 0, sizeof(*stats));
  int s = krylov_solve_noalloc(*ctx, A, r, x, *opts, stats);
  MHX_HIP(hipStreamSynchronize(ctx->stream));
  if (stats) stats->status = s;
  return s;
  MHX_GUARD_END(ctx)
}

// nodal history terms of a time step / Butcher stage
// (Workset::computeSolnTransientSeeded, workset.cpp:601-680):
//   beta_t = (sum_{k>=1} BDF_k u_prev[k-1]) * timewt
//   beta_u = (1 - alpha_u) u_prev[0] + sum_{j<s} A_sj / b_j (U_j - u_prev[0])
// in the reference's order of operations
constexpr int kMaxStages = 8;
struct StageCoef {
  double w[7];            // BDF_1..BDF_k
  double a[kMaxStages];   // A_sj / b_j, j < s
};
__global__ void k_stage_history(long n, int order, const double* __restrict__ prev, int stage,
                                const double* __restrict__ ustage, double timewt, double alpha_u, StageCoef sc,
                                double* __restrict__ beta_t, double* __restrict__ beta_u) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double bt = 0.0;
  for (int k = 1; k <= order; ++k) bt += sc.w[k - 1] * prev[(long)(k - 1) * n + i];
  beta_t[i] = bt * timewt;
  if (beta_u) {
    const double u0 = prev[i];
    double bu = (1.0 - alpha_u) * u0;
    for (int j = 0; j < stage; ++j) bu += sc.a[j] * (ustage[(long)j * n + i] - u0);
    beta_u[i] = bu;
  }
}

static void set_stage_state(Ctx& c, double dt, int32_t bdf_order, const double* u_prev, int32_t n_stages,
                            const double* A, const double* b, int32_t stage, const double* u_stages, int32_t flags) {
  MHX_REQUIRE(bdf_order >= 0 && bdf_order <= 6, MHX_ERR_ARG, "bdf_order must be 0..6");
  if (bdf_order == 0) {
    c.bdf_order = 0;
    c.alpha_t = 0.0;
    c.alpha_u = 1.0;
    c.hist_u.release();
    return;
  }
  MHX_REQUIRE(c.desc.physics == MHX_PHYS_THERMAL, MHX_ERR_UNSUPPORTED, "time integration is built for the thermal module");
  MHX_REQUIRE(dt > 0.0 && std::isfinite(dt), MHX_ERR_ARG, "dt must be positive");
  MHX_REQUIRE(u_prev, MHX_ERR_ARG, "null history");
  MHX_REQUIRE(n_stages >= 1 && n_stages <= kMaxStages && stage >= 0 && stage < n_stages && A && b, MHX_ERR_ARG,
              "Butcher tableau: 1..8 stages, stage in range, A and b given");
  MHX_REQUIRE(stage == 0 || u_stages, MHX_ERR_ARG, "the earlier stages' solutions are needed");
  const double bs = b[stage], ass = A[stage * n_stages + stage];
  MHX_REQUIRE(bs != 0.0 && std::isfinite(bs) && std::isfinite(ass), MHX_ERR_ARG, "Butcher b of the stage must be nonzero");
  for (int j = 0; j < stage; ++j) MHX_REQUIRE(b[j] != 0.0, MHX_ERR_ARG, "Butcher b of an earlier stage is zero");
  const bool multi = n_stages > 1 || ass != 1.0 || bs != 1.0;
  MHX_REQUIRE(!multi || !c.desc.wdbc_sides, MHX_ERR_UNSUPPORTED,
              "weak Dirichlet sides take the stage unknown; built for one-stage tableaus only");
  // backward-difference weights without 1/dt (solverManager_setup.hpp:455-500)
  static const double W[7][7] = {
      {0}, {1.0, -1.0}, {1.5, -2.0, 0.5}, {11.0 / 6.0, -3.0, 1.5, -1.0 / 3.0},
      {25.0 / 12.0, -4.0, 3.0, -4.0 / 3.0, 1.0 / 4.0},
      {137.0 / 60.0, -5.0, 5.0, -10.0 / 3.0, 75.0 / 60.0, -1.0 / 5.0},
      {147.0 / 60.0, -6.0, 15.0 / 2.0, -20.0 / 3.0, 225.0 / 60.0, -72.0 / 60.0, 1.0 / 6.0}};
  const long n = c.mesh.ndof;
  const bool dev = (flags & MHX_DEVICE_PTRS) != 0;
  const double* prev = u_prev;
  const double* ust = u_stages;
  DevBuf<double> stg_prev, stg_st;
  if (!dev) {
    stg_prev.alloc((size_t)n * bdf_order);
    MHX_HIP(hipMemcpyAsync(stg_prev.p, u_prev, (size_t)n * bdf_order * sizeof(double), hipMemcpyHostToDevice,
                           c.stream));
    prev = stg_prev.p;
    if (stage > 0) {
      stg_st.alloc((size_t)n * stage);
      MHX_HIP(hipMemcpyAsync(stg_st.p, u_stages, (size_t)n * stage * sizeof(double), hipMemcpyHostToDevice,
                             c.stream));
      ust = stg_st.p;
    }
  }
  const double timewt = 1.0 / dt / bs;
  const double alpha_u = ass / bs;
  StageCoef sc{};
  for (int k = 1; k <= bdf_order; ++k) sc.w[k - 1] = W[bdf_order][k];
  for (int j = 0; j < stage; ++j) sc.a[j] = A[stage * n_stages + j] / b[j];
  c.hist.alloc(n);
  if (multi) c.hist_u.alloc(n);
  else c.hist_u.release();
  hipLaunchKernelGGL(k_stage_history, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, n, (int)bdf_order,
                     prev, (int)stage, ust, timewt, alpha_u, sc, c.hist.p, multi ? c.hist_u.p : nullptr);
  MHX_HIP(hipGetLastError());
  c.alpha_t = W[bdf_order][0] * timewt;
  c.alpha_u = multi ? alpha_u : 1.0;
  c.bdf_order = bdf_order;
  c.jac_valid = false;
  MHX_HIP(hipStreamSynchronize(c.stream));
}

int mhx_set_time_state(mhx_ctx* ctx, double dt, int32_t bdf_order, const double* u_prev, int32_t flags) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx, MHX_ERR_ARG, "null argument");
  const double one = 1.0;
  set_stage_state(*ctx, dt, bdf_order, u_prev, 1, &one, &one, 0, nullptr, flags);
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_set_stage_state(mhx_ctx* ctx, double dt, int32_t bdf_order, const double* u_prev, int32_t n_stages,
                        const double* butcher_A, const double* butcher_b, int32_t stage, const double* u_stages,
                        int32_t flags) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx, MHX_ERR_ARG, "null argument");
  set_stage_state(*ctx, dt, bdf_order, u_prev, n_stages, butcher_A, butcher_b, stage, u_stages, flags);
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_write_matrix_market(mhx_ctx* ctx, const char* path, const double* jac_vals, int32_t flags) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && path, MHX_ERR_ARG, "null argument");
  bool dev = (flags & MHX_DEVICE_PTRS) != 0;
  if (!jac_vals) {
    MHX_REQUIRE(ctx->jac_valid && ctx->jac.p, MHX_ERR_STATE, "no Jacobian held by the context");
    jac_vals = ctx->jac.p;
    dev = true;
  }
  write_matrix_market(*ctx, path, jac_vals, dev);
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_write_vector_market(mhx_ctx* ctx, const char* path, const double* vec, int32_t flags) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && path && vec, MHX_ERR_ARG, "null argument");
  write_vector_market(*ctx, path, vec, (flags & MHX_DEVICE_PTRS) != 0);
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_read_matrix_market(mhx_ctx* ctx, const char* path, double* jac_vals, int32_t flags) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && path && jac_vals, MHX_ERR_ARG, "null argument");
  read_matrix_market(*ctx, path, jac_vals, (flags & MHX_DEVICE_PTRS) != 0);
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_read_vector_market(mhx_ctx* ctx, const char* path, double* vec, int32_t flags) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && path && vec, MHX_ERR_ARG, "null argument");
  read_vector_market(*ctx, path, vec, (flags & MHX_DEVICE_PTRS) != 0);
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_amg_setup(mhx_ctx* ctx, const double* jac_vals, int32_t flags) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx, MHX_ERR_ARG, "null argument");
  const double* A;
  DevBuf<double> jstage;
  if (jac_vals) {
    if (flags & MHX_DEVICE_PTRS) {
      A = jac_vals;
    } else {
      jstage.alloc(ctx->nnz);
      MHX_HIP(hipMemcpyAsync(jstage.p, jac_vals, ctx->nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      A = jstage.p;
    }
  } else {
    MHX_REQUIRE(ctx->jac_valid && ctx->jac.p, MHX_ERR_STATE, "no Jacobian held by the context");
    A = ctx->jac.p;
  }
  amg_setup(*ctx, A, false);
  MHX_HIP(hipStreamSynchronize(ctx->stream));
  return amg_levels(*ctx);
  MHX_GUARD_END(ctx)
}

int mhx_amg_level(mhx_ctx* ctx, int32_t level, int32_t which, int64_t info[4], double* lambda, int32_t* rp,
                  int32_t* ci, double* vals) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && info, MHX_ERR_ARG, "null argument");
  MHX_REQUIRE(amg_levels(*ctx) > 0, MHX_ERR_STATE, "no AMG hierarchy held by the context");
  MHX_REQUIRE(!(level == 0 && which == 0 && vals), MHX_ERR_ARG,
              "the level-0 operator is the caller's Jacobian; its values are not kept");
  amg_download(*ctx, level, which, &info[0], &info[1], &info[2], rp, ci, vals);
  info[3] = amg_levels(*ctx);
  if (lambda) *lambda = amg_lambda(*ctx, level);
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_nonlinear_solve(mhx_ctx* ctx, double* u, const mhx_newton_opts* opts, mhx_solve_stats* stats) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && u && opts, MHX_ERR_ARG, "null argument");
  const bool dev = (opts->flags & MHX_DEVICE_PTRS) != 0;
  const long n = ctx->mesh.ndof;
  VecArg U(*ctx, u, n, dev, ctx->stage_u, true);
  if (stats) std::memset(stats, 0, sizeof(*stats));
  int s = newton(*ctx, U.dev, *opts, stats);
  U.copy_out(u);
  MHX_HIP(hipStreamSynchronize(ctx->stream));
  if (stats) stats->status = s;
  return s;
  MHX_GUARD_END(ctx)
}

int mhx_l2_error(mhx_ctx* ctx, const double* u, int32_t n_modes, const double* amp, const double* freq, int32_t flags,
                 double* l2, double* l2_grad) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && u && l2 && l2_grad, MHX_ERR_ARG, "null argument");
  MHX_REQUIRE(n_modes >= 0 && n_modes <= MHX_MAX_MODES, MHX_ERR_ARG, "n_modes out of range");
  MHX_REQUIRE(ctx->mesh.nf == 1, MHX_ERR_UNSUPPORTED, "L2 error of a single-field (thermal) solution only");
  const bool dev = (flags & MHX_DEVICE_PTRS) != 0;
  VecArg U(*ctx, u, ctx->mesh.ndof, dev, ctx->stage_a, true);
  ErrArgs a{};
  a.m = ctx->mesh;
  a.t = ctx->tab;
  a.vx = ctx->vx.p;
  a.u = U.dev;
  a.nm = n_modes;
  for (int m = 0; m < n_modes; ++m) {
    a.amp[m] = amp[m];
    for (int d = 0; d < 3; ++d) a.freq[m][d] = freq[m * 3 + d];
  }
  const int nb = 256;
  DevBuf<double> part;
  part.alloc(2 * nb);
  a.part = part.p;
  hipLaunchKernelGGL(k_l2err, dim3(nb), dim3(256), 0, ctx->stream, a);
  MHX_HIP(hipGetLastError());
  std::vector<double> h(2 * nb);
  MHX_HIP(hipMemcpyAsync(h.data(), part.p, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MHX_HIP(hipStreamSynchronize(ctx->stream));
  double s0 = 0.0, s1 = 0.0;
  for (int i = 0; i < nb; ++i) {
    s0 += h[2 * i];
    s1 += h[2 * i + 1];
  }
  *l2 = std::sqrt(s0);
  *l2_grad = std::sqrt(s1);
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_profile_enable(mhx_ctx* ctx, int32_t on) {
  if (!ctx) return MHX_ERR_ARG;
  ctx->prof.on = on != 0;
  return MHX_OK;
}

int mhx_profile_read(mhx_ctx* ctx, int32_t cap, char* names48, double* ms, int64_t* counts) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx, MHX_ERR_ARG, "null ctx");
  MHX_HIP(hipStreamSynchronize(ctx->stream));
  ctx->prof.drain();
  const int n = (int)ctx->prof.phases.size();
  for (int i = 0; i < n && i < cap; ++i) {
    if (names48) {
      std::memset(names48 + 48 * i, 0, 48);
      std::strncpy(names48 + 48 * i, ctx->prof.phases[i].name.c_str(), 47);
    }
    if (ms) ms[i] = ctx->prof.phases[i].ms;
    if (counts) counts[i] = ctx->prof.phases[i].count;
  }
  return n;
  MHX_GUARD_END(ctx)
}

int mhx_profile_reset(mhx_ctx* ctx) {
  if (!ctx) return MHX_ERR_ARG;
  (void)hipStreamSynchronize(ctx->stream);
  ctx->prof.drain();
  for (auto& p : ctx->prof.phases) {
    p.ms = 0.0;
    p.count = 0;
  }
  return MHX_OK;
}

int mhx_synchronize(mhx_ctx* ctx) {
  if (!ctx) return MHX_ERR_ARG;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    ctx->err = hipGetErrorString(e);
    return MHX_ERR_DEVICE;
  }
  return MHX_OK;
}

// sum over the lattice nodes of one axis of the 1-D row length (the CRS graph
// is the tensor product of these ranges)
static int64_t axis_pattern_sum(int p, int nel) {
  const int nn = p * nel + 1;
  int64_t s = 0;
  for (int c = 0; c < nn; ++c) {
    if (c % p == 0) s += std::min(c + p, nn - 1) - std::max(c - p, 0) + 1;
    else s += p + 1;
  }
  return s;
}

int mhx_footprint(const mhx_problem_desc* d, const mhx_krylov_opts* opts, int64_t* n_dof, int64_t* nnz,
                  int64_t* device_bytes) {
  MHX_GUARD_BEGIN
  validate(d);
  const int nf = d->physics == MHX_PHYS_NAVIER_STOKES ? d->dim + 1 : 1;
  const int nel[3] = {d->nx, d->ny, d->dim == 3 ? d->nz : 1};
  int64_t N = nf, Z = (int64_t)nf * nf, ne = 1, nv = 1;
  for (int a = 0; a < d->dim; ++a) {
    N *= (int64_t)d->order * nel[a] + 1;
    Z *= axis_pattern_sum(d->order, nel[a]);
    ne *= nel[a];
    nv *= nel[a] + 1;
  }
  // CRS graph + values; u, res, du, aux, staging, Dirichlet tables; CG
  // vectors r, z, p, q and the preconditioner; mesh arrays; reduction space
  int64_t bytes = 12 * Z + 4 * (N + 1) + N * (8 * 4 + 8 * 3 + 8 + 1 + 4) + 8 * 5 * N + 24 * nv + 16 * ne + (1 << 20);
  if (opts && opts->solver != MHX_KRYLOV_CG) bytes += 8 * N * ((opts->restart > 0 ? opts->restart : 50) + 3);
  if (opts && (opts->precond & MHX_PREC_FP32)) bytes += 4 * Z + 4 * 8 * N;
  if (opts && (opts->precond & 0xff) == MHX_PREC_AMG) {
    // fine level: A P in 27-slot boxes per (row, field) during set-up, P and
    // R (at most 8 nf entries per fine row each), five level vectors and the
    // power-iteration pair; coarse levels shrink by 27x or more each
    const int64_t fine = 8 * 27 * N * nf + 2 * 12 * 8 * N * nf + 8 * 7 * N;
    bytes += fine + fine / 8;
  }
  if (n_dof) *n_dof = N;
  if (nnz) *nnz = Z;
  if (device_bytes) *device_bytes = bytes;
  return MHX_OK;
  MHX_GUARD_END((mhx_ctx*)nullptr)
}

int mhx_ctx_update(mhx_ctx* ctx, const mhx_problem_desc* d) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && d, MHX_ERR_ARG, "null argument");
  validate(d);
  const mhx_problem_desc& o = ctx->desc;
  const bool same = d->dim == o.dim && d->order == o.order && d->quadrature == o.quadrature && d->nx == o.nx &&
                    d->ny == o.ny && (d->dim == 2 || d->nz == o.nz) && d->xmin == o.xmin && d->xmax == o.xmax &&
                    d->ymin == o.ymin && d->ymax == o.ymax && d->zmin == o.zmin && d->zmax == o.zmax &&
                    d->perturb == o.perturb && d->perturb_seed == o.perturb_seed && d->physics == o.physics &&
                    d->dbc_sides == o.dbc_sides && d->ns_lid_sides == o.ns_lid_sides &&
                    d->ns_pin_pressure == o.ns_pin_pressure && d->ns_comp_sides[0] == o.ns_comp_sides[0] &&
                    d->ns_comp_sides[1] == o.ns_comp_sides[1] && d->ns_comp_sides[2] == o.ns_comp_sides[2];
  MHX_REQUIRE(same, MHX_ERR_ARG, "ctx_update: mesh, discretisation or Dirichlet sides differ from the context's");
  Ctx& c = *ctx;
  const BoxMesh& m = c.mesh;
  auto upload = [&](const double* src, DevBuf<double>& buf) {
    if (src) {
      buf.alloc(m.ne);
      MHX_HIP(hipMemcpyAsync(buf.p, src, m.ne * sizeof(double), hipMemcpyHostToDevice, c.stream));
    } else {
      buf.release();
    }
  };
  upload(d->k0, c.k0);
  upload(d->src_elem, c.src);
  // coefficient functions (new expressions or a new time)
  ExprHost xh;
  expr_compile(*d, xh);
  MHX_REQUIRE(!xh.active || d->physics == MHX_PHYS_THERMAL, MHX_ERR_UNSUPPORTED,
              "coefficient functions are built for the thermal module");
  MHX_REQUIRE(!xh.active || !d->wdbc_sides, MHX_ERR_UNSUPPORTED,
              "weak Dirichlet sides take the closed-form diffusion; not built with coefficient functions");
  c.xh = std::move(xh);
  upload_functions(c);
  if (c.xh.active && c.scatter_mode == kPencilStream) c.scatter_mode = kElementColouring;
  // Dirichlet values on the (unchanged) Dirichlet DOFs
  const bool ns = d->physics == MHX_PHYS_NAVIER_STOKES;
  for (int32_t g : c.h_dbc_list) {
    if (!ns) {
      c.h_dbc[g] = d->dbc_values ? d->dbc_values[g] : d->dbc_const;
    } else {
      const int f = g % m.nf;
      const long node = g / m.nf;
      const long gc[3] = {node % m.nn[0], (node / m.nn[0]) % m.nn[1], node / ((long)m.nn[0] * m.nn[1])};
      uint32_t on = 0;
      for (int a = 0; a < m.dim; ++a) {
        if (gc[a] == 0) on |= 1u << (2 * a);
        if (gc[a] == m.nn[a] - 1) on |= 1u << (2 * a + 1);
      }
      c.h_dbc[g] = (f < m.dim && (on & d->ns_lid_sides)) ? d->ns_lid_u[f] : 0.0;
    }
  }
  MHX_HIP(hipMemcpyAsync(c.dbc_vals.p, c.h_dbc.data(), m.ndof * sizeof(double), hipMemcpyHostToDevice, c.stream));
  MHX_HIP(hipStreamSynchronize(c.stream));
  c.desc = *d;
  c.desc.k0 = nullptr;
  c.desc.src_elem = nullptr;
  c.desc.dbc_values = nullptr;
  c.desc.function_names = nullptr;
  c.desc.function_exprs = nullptr;
  c.jac_valid = false;
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_set_option(mhx_ctx* ctx, const char* name, int64_t value) {
  if (!ctx || !name) return MHX_ERR_ARG;
  const std::string k(name);
  if (k == "assembly.colouring") {
    if (value != kElementColouring && value != kBlockColouring && value != kSegmentMerge &&
        !(value == kPencilStream && q2box_applicable(*ctx))) {
      ctx->err = "assembly.colouring must be 0, 1 or 2 (3: Q2 thermal on a box mesh only)";
      return MHX_ERR_ARG;
    }
    ctx->scatter_mode = (int)value;
    return MHX_OK;
  }
  if (k == "assembly.segment") {
    if (value < 2 || value > 4096 || (value & 1)) {
      ctx->err = "assembly.segment must be an even element count in [2, 4096]";
      return MHX_ERR_ARG;
    }
    ctx->q2_segment = (int)value;
    return MHX_OK;
  }
  if (k == "krylov.spmv_lanes") {
    if (value != 4 && value != 8 && value != 16 && value != 32 && value != 64) {
      ctx->err = "krylov.spmv_lanes must be 4, 8, 16, 32 or 64";
      return MHX_ERR_ARG;
    }
    ctx->spmv_lanes = (int)value;
    return MHX_OK;
  }
  if (k == "krylov.lattice_columns") {
    if (value != 0 && value != 1) {
      ctx->err = "krylov.lattice_columns must be 0 or 1";
      return MHX_ERR_ARG;
    }
    ctx->lattice_cols = (int)value;
    return MHX_OK;
  }
  if (k == "amg.reuse") {
    if (value != 0 && value != 1) {
      ctx->err = "amg.reuse must be 0 or 1";
      return MHX_ERR_ARG;
    }
    ctx->amg_reuse = (int)value;
    return MHX_OK;
  }
  if (k == "amg.sweeps") {
    if (value < 1 || value > 16) {
      ctx->err = "amg.sweeps must be 1 .. 16";
      return MHX_ERR_ARG;
    }
    ctx->amg_sweeps = (int)value;
    return MHX_OK;
  }
  if (k == "amg.omega_pct") {
    if (value < 1 || value > 200) {
      ctx->err = "amg.omega_pct must be 1 .. 200";
      return MHX_ERR_ARG;
    }
    ctx->amg_omega_pct = (int)value;
    return MHX_OK;
  }
  if (k == "amg.smoother") {
    if (value < -1 || value > 2) {
      ctx->err = "amg.smoother must be -1 (automatic), 0 (Chebyshev), 1 (nodal block Jacobi) or 2 (nodal block Gauss-Seidel)";
      return MHX_ERR_ARG;
    }
    ctx->amg_smoother = (int)value;
    return MHX_OK;
  }
  ctx->err = "unknown option " + k;
  return MHX_ERR_ARG;
}

int mhx_get_option(const mhx_ctx* ctx, const char* name, int64_t* value) {
  if (!ctx || !name || !value) return MHX_ERR_ARG;
  const std::string k(name);
  if (k == "assembly.segment") {
    *value = ctx->q2_segment;
    return MHX_OK;
  }
  if (k == "assembly.colouring") {
    *value = ctx->scatter_mode;
    return MHX_OK;
  }
  if (k == "krylov.spmv_lanes") {
    *value = ctx->spmv_lanes;
    return MHX_OK;
  }
  if (k == "krylov.lattice_columns") {
    *value = ctx->lattice_cols;
    return MHX_OK;
  }
  if (k == "amg.reuse") {
    *value = ctx->amg_reuse;
    return MHX_OK;
  }
  if (k == "amg.sweeps") {
    *value = ctx->amg_sweeps;
    return MHX_OK;
  }
  if (k == "amg.omega_pct") {
    *value = ctx->amg_omega_pct;
    return MHX_OK;
  }
  if (k == "amg.smoother") {
    *value = ctx->amg_smoother;
    return MHX_OK;
  }
  return MHX_ERR_ARG;
}

static DevBuf<double>& state_buf(mhx_ctx* ctx, int32_t which, size_t& count) {
  switch (which) {
    case MHX_BUF_U: count = ctx->mesh.ndof; return ctx->u;
    case MHX_BUF_RES: count = ctx->mesh.ndof; return ctx->res;
    case MHX_BUF_JAC: count = ctx->nnz; return ctx->jac;
    case MHX_BUF_DU: count = ctx->mesh.ndof; return ctx->du;
    case MHX_BUF_AUX: count = ctx->mesh.ndof; return ctx->aux;
    default: throw Error{MHX_ERR_ARG, "unknown state buffer"};
  }
}

int mhx_state_ptr(mhx_ctx* ctx, int32_t which, double** dev_ptr, int64_t* count) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && dev_ptr, MHX_ERR_ARG, "null argument");
  size_t n = 0;
  DevBuf<double>& b = state_buf(ctx, which, n);
  if (b.n != n) {
    b.alloc(n);
    MHX_HIP(hipMemsetAsync(b.p, 0, n * sizeof(double), ctx->stream));
    MHX_HIP(hipStreamSynchronize(ctx->stream));
  }
  *dev_ptr = b.p;
  if (count) *count = (int64_t)n;
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_state_upload(mhx_ctx* ctx, int32_t which, const double* host, int64_t count) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && host, MHX_ERR_ARG, "null argument");
  size_t n = 0;
  DevBuf<double>& b = state_buf(ctx, which, n);
  MHX_REQUIRE((size_t)count == n, MHX_ERR_ARG, "count does not match the buffer size");
  if (b.n != n) b.alloc(n);
  MHX_HIP(hipMemcpyAsync(b.p, host, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  MHX_HIP(hipStreamSynchronize(ctx->stream));
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_state_download(mhx_ctx* ctx, int32_t which, double* host, int64_t count) {
  MHX_GUARD_BEGIN
  MHX_REQUIRE(ctx && host, MHX_ERR_ARG, "null argument");
  size_t n = 0;
  DevBuf<double>& b = state_buf(ctx, which, n);
  MHX_REQUIRE((size_t)count == n, MHX_ERR_ARG, "count does not match the buffer size");
  MHX_REQUIRE(b.n == n && b.p, MHX_ERR_STATE, "buffer not allocated");
  MHX_HIP(hipMemcpyAsync(host, b.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MHX_HIP(hipStreamSynchronize(ctx->stream));
  return MHX_OK;
  MHX_GUARD_END(ctx)
}

int mhx_set_stream(mhx_ctx* ctx, void* hip_stream) {
  if (!ctx) return MHX_ERR_ARG;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    ctx->err = hipGetErrorString(e);
    return MHX_ERR_DEVICE;
  }
  if (ctx->own_stream && ctx->stream != (hipStream_t)hip_stream) {
    (void)hipStreamDestroy(ctx->stream);
    ctx->own_stream = false;
  }
  ctx->stream = (hipStream_t)hip_stream;
  return MHX_OK;
}

int mhx_copy_probe(int64_t n_doubles, int32_t reps, double* ms_per_copy) {
  try {
    if (n_doubles <= 0 || reps <= 0 || !ms_per_copy) return MHX_ERR_ARG;
    DevBuf<double> x, y;
    x.alloc(n_doubles);
    y.alloc(n_doubles);
    hipStream_t s;
    MHX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    MHX_HIP(hipMemsetAsync(x.p, 0, n_doubles * sizeof(double), s));
    const unsigned grid = (unsigned)std::min<int64_t>((n_doubles / 2 + 255) / 256, 65536);
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s, n_doubles, x.p, y.p);
    hipEvent_t a, b;
    MHX_HIP(hipEventCreate(&a));
    MHX_HIP(hipEventCreate(&b));
    MHX_HIP(hipEventRecord(a, s));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s, n_doubles, x.p, y.p);
    MHX_HIP(hipEventRecord(b, s));
    MHX_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    MHX_HIP(hipEventElapsedTime(&ms, a, b));
    *ms_per_copy = ms / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipStreamDestroy(s);
    return MHX_OK;
  } catch (const Error& e) {
    g_last_error = e.msg;
    return e.code;
  }
}

static inline uint64_t splitmix(uint64_t seed, uint64_t idx) {
  uint64_t z = seed ^ (idx * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int mhx_rng_fill(uint64_t seed, int64_t n, int32_t kind, double a, double b, double* out) {
  if (n < 0 || (n && !out)) return MHX_ERR_ARG;
  for (int64_t i = 0; i < n; ++i) {
    double u1 = (double)(splitmix(seed, 2 * (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0);
    if (kind == MHX_RNG_UNIFORM) {
      out[i] = a + (b - a) * u1;
    } else {
      double u2 = (double)(splitmix(seed, 2 * (uint64_t)i + 1) >> 11) * (1.0 / 9007199254740992.0);
      double z = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
      out[i] = kind == MHX_RNG_NORMAL ? a + b * z : std::exp(a + b * z);
    }
  }
  return MHX_OK;
}

}  // extern "C"
