// krylov.hip — CRS SpMV, wavefront-reduced dot products and the Jacobi-
// preconditioned CG / restarted GMRES(m) drivers (FP64, optional FP32 inner
// solve with FP64 iterative refinement).
//
// Replaces LinearAlgebraInterface::linearSolver (Belos Block GMRES / CG with
// Ifpack2 RELAXATION Jacobi; linearAlgebraInterface_solvers.hpp:14-152) and
// SolverManager::PCG (solverManager_solvers.hpp:812-905).
// Stopping rule for both: ||b - A x||_2 / ||b||_2 <= tol, FP64.

#include <algorithm>
#include <cmath>
#include <vector>

#include "mhx_internal.hpp"

namespace mhx {

constexpr int kRedBlocks = 1024;  // partial sums per reduction

// y = A (s .* x)  (s == nullptr: plain SpMV).  LANES lanes per row.
template <int LANES, class TA, class TX>
__global__ __launch_bounds__(256) void k_spmv(long n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                              const TA* __restrict__ A, const TX* __restrict__ s,
                                              const TX* __restrict__ x, TX* __restrict__ y) {
  const long row = ((long)blockIdx.x * blockDim.x + threadIdx.x) / LANES;
  const int l = threadIdx.x % LANES;
  if (row >= n) return;
  const int b = rp[row], e = rp[row + 1];
  TX acc = 0;
  if (s) {
    for (int k = b + l; k < e; k += LANES) {
      const int c = ci[k];
      acc += (TX)A[k] * (s[c] * x[c]);
    }
  } else {
    for (int k = b + l; k < e; k += LANES) acc += (TX)A[k] * x[ci[k]];
  }
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
  if (l == 0) y[row] = acc;
}

template <class TA, class TX>
static void spmv_t(Ctx& c, const TA* A, const TX* s, const TX* x, TX* y) {
  long n = c.mesh.ndof;
  int L = c.spmv_lanes;
  long threads = n * L;
  dim3 grid((unsigned)((threads + 255) / 256));
  switch (L) {
    case 4: hipLaunchKernelGGL((k_spmv<4, TA, TX>), grid, dim3(256), 0, c.stream, n, c.rowptr.p, c.colidx.p, A, s, x, y); break;
    case 8: hipLaunchKernelGGL((k_spmv<8, TA, TX>), grid, dim3(256), 0, c.stream, n, c.rowptr.p, c.colidx.p, A, s, x, y); break;
    case 16: hipLaunchKernelGGL((k_spmv<16, TA, TX>), grid, dim3(256), 0, c.stream, n, c.rowptr.p, c.colidx.p, A, s, x, y); break;
    case 32: hipLaunchKernelGGL((k_spmv<32, TA, TX>), grid, dim3(256), 0, c.stream, n, c.rowptr.p, c.colidx.p, A, s, x, y); break;
    default: hipLaunchKernelGGL((k_spmv<64, TA, TX>), grid, dim3(256), 0, c.stream, n, c.rowptr.p, c.colidx.p, A, s, x, y); break;
  }
  MHX_HIP(hipGetLastError());
}

void spmv(Ctx& c, const double* A, const double* x, double* y) {
  ScopedPhase ph(c.prof, "Krylov::spmv", c.stream);
  spmv_t<double, double>(c, A, nullptr, x, y);
}

// ---- reductions: per-block partials in fixed order (deterministic) ----------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int K>
__device__ __forceinline__ void block_store(double (&v)[K], double* out, int stride) {
  __shared__ double sh[K][4];
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = wave_sum(v[k]);
    if (l == 0) sh[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    out[(long)k * stride + blockIdx.x] = sh[k][0] + sh[k][1] + sh[k][2] + sh[k][3];
  }
}

// max that propagates NaN (fmax would drop it and hide a non-finite residual)
__device__ __forceinline__ double nanmax(double v, double x) { return (v != v || x <= v) ? v : x; }

// mode 0: sum a*b ; mode 1: max |a| (NaN-propagating)
__global__ __launch_bounds__(256) void k_dot(long n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* __restrict__ part, int mode) {
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    if (mode == 0) v[0] = fma(a[i], b[i], v[0]);
    else v[0] = nanmax(v[0], fabs(a[i]));
  }
  if (mode == 0) {
    block_store<1>(v, part, kRedBlocks);
  } else {
    __shared__ double sh[256];
    sh[threadIdx.x] = v[0];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] = nanmax(sh[threadIdx.x], sh[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
  }
}

// reduce K rows of kRedBlocks partials (row stride kRedBlocks) into out[K]
__global__ __launch_bounds__(256) void k_final(const double* __restrict__ part, int nb, int K, int mode,
                                               double* __restrict__ out) {
  for (int k = 0; k < K; ++k) {
    double v[1] = {0.0};
    for (int i = threadIdx.x; i < nb; i += 256) {
      double p = part[(long)k * kRedBlocks + i];
      v[0] = mode == 0 ? v[0] + p : nanmax(v[0], p);
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = v[0];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] = mode == 0 ? sh[threadIdx.x] + sh[threadIdx.x + s] : nanmax(sh[threadIdx.x], sh[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sh[0];
    __syncthreads();
  }
}

static int red_blocks(long n) {
  long b = (n + 255) / 256;
  return (int)std::min<long>(std::max<long>(b, 1), kRedBlocks);
}

static void ensure_red(Ctx& c, int K) {
  if (c.kw.part.n < (size_t)K * kRedBlocks) c.kw.part.alloc((size_t)K * kRedBlocks);
  if (c.kw.scal.n < (size_t)std::max(K + 160, 256)) c.kw.scal.alloc(std::max(K + 160, 256));
}

static double reduce_to_host(Ctx& c, long n, const double* a, const double* b, int mode) {
  ensure_red(c, 1);
  int nb = red_blocks(n);
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(256), 0, c.stream, n, a, b, c.kw.part.p, mode);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, c.stream, c.kw.part.p, nb, 1, mode, c.kw.scal.p);
  MHX_HIP(hipGetLastError());
  double h;
  MHX_HIP(hipMemcpyAsync(&h, c.kw.scal.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MHX_HIP(hipStreamSynchronize(c.stream));
  return h;
}

double dot(Ctx& c, const double* a, const double* b, long n) {
  ScopedPhase ph(c.prof, "Krylov::dot", c.stream);
  return reduce_to_host(c, n, a, b, 0);
}
double norm2(Ctx& c, const double* a, long n) { return std::sqrt(dot(c, a, a, n)); }
double norm_inf(Ctx& c, const double* a, long n) {
  ScopedPhase ph(c.prof, "Krylov::normInf", c.stream);
  return reduce_to_host(c, n, a, a, 1);
}

// ---- elementwise kernels ----------------------------------------------------
__global__ void k_axpy(long n, double a, const double* __restrict__ x, double* __restrict__ y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = fma(a, x[i], y[i]);
}
void axpy(Ctx& c, long n, double a, const double* x, double* y) {
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, n, a, x, y);
  MHX_HIP(hipGetLastError());
}

// minv[i] = 1 / A(i,i)
template <class T>
__global__ void k_jacobi(long n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                         const double* __restrict__ A, T* __restrict__ minv) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = rp[i], hi = rp[i + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ci[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  minv[i] = (T)(1.0 / A[lo]);
}

__global__ void k_set_dbc(long n, const int32_t* __restrict__ list, const double* __restrict__ vals,
                          double* __restrict__ u) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[list[i]] = vals[list[i]];
}
void set_dirichlet_dev(Ctx& c, double* u) {
  long n = (long)c.h_dbc_list.size();
  if (!n) return;
  hipLaunchKernelGGL(k_set_dbc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, n, c.dbc_list.p, c.dbc_vals.p, u);
  MHX_HIP(hipGetLastError());
}

// r = b - A x ; returns ||r||^2 partials path via dot afterwards
__global__ void k_resid(long n, const double* __restrict__ b, const double* __restrict__ Ax, double* __restrict__ r) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - Ax[i];
}

// ---- CG (Jacobi PCG as SolverManager::PCG) -----------------------------------
// z = minv .* r  fused with rho = r.z
__global__ __launch_bounds__(256) void k_prec_dot(long n, const double* __restrict__ minv, const double* __restrict__ r,
                                                  double* __restrict__ z, double* __restrict__ part) {
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    double zi = minv[i] * r[i];
    z[i] = zi;
    v[0] = fma(r[i], zi, v[0]);
  }
  block_store<1>(v, part, kRedBlocks);
}
// p = z + beta p
__global__ void k_xpby(long n, const double* __restrict__ z, double beta, double* __restrict__ p) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = fma(beta, p[i], z[i]);
}
// x += alpha p ; r -= alpha q ; partial ||r||^2
__global__ __launch_bounds__(256) void k_cg_update(long n, double alpha, const double* __restrict__ p,
                                                   const double* __restrict__ q, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ part) {
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    x[i] = fma(alpha, p[i], x[i]);
    double ri = fma(-alpha, q[i], r[i]);
    r[i] = ri;
    v[0] = fma(ri, ri, v[0]);
  }
  block_store<1>(v, part, kRedBlocks);
}

static double fetch_scalar(Ctx& c, int nb) {
  hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, c.stream, c.kw.part.p, nb, 1, 0, c.kw.scal.p);
  double h;
  MHX_HIP(hipMemcpyAsync(&h, c.kw.scal.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MHX_HIP(hipStreamSynchronize(c.stream));
  return h;
}

static int cg_solve(Ctx& c, const double* A, const double* b, double* x, double tol, int maxit, double* relres) {
  const long n = c.mesh.ndof;
  KrylovWork& w = c.kw;
  w.r.alloc(n); w.z.alloc(n); w.p.alloc(n); w.q.alloc(n); w.minv.alloc(n);
  ensure_red(c, 1);
  const dim3 g1((unsigned)((n + 255) / 256));
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_jacobi<double>, g1, dim3(256), 0, c.stream, n, c.rowptr.p, c.colidx.p, A, w.minv.p);
  spmv(c, A, x, w.q.p);
  hipLaunchKernelGGL(k_resid, g1, dim3(256), 0, c.stream, n, b, w.q.p, w.r.p);
  const double bn = norm2(c, b, n);
  double rn = norm2(c, w.r.p, n);
  if (bn == 0.0) {
    MHX_HIP(hipMemsetAsync(x, 0, n * sizeof(double), c.stream));
    *relres = 0.0;
    return 0;
  }
  double rho = 1.0;
  int it = 0;
  while (it < maxit && rn / bn > tol) {
    ScopedPhase ph(c.prof, "Krylov::cg_iteration", c.stream);
    hipLaunchKernelGGL(k_prec_dot, dim3(nb), dim3(256), 0, c.stream, n, w.minv.p, w.r.p, w.z.p, w.part.p);
    const double rho1 = rho;
    rho = fetch_scalar(c, nb);
    const double beta = it == 0 ? 0.0 : rho / rho1;
    hipLaunchKernelGGL(k_xpby, g1, dim3(256), 0, c.stream, n, w.z.p, beta, w.p.p);
    spmv_t<double, double>(c, A, nullptr, w.p.p, w.q.p);
    const double pq = reduce_to_host(c, n, w.p.p, w.q.p, 0);
    const double alpha = rho / pq;
    hipLaunchKernelGGL(k_cg_update, dim3(nb), dim3(256), 0, c.stream, n, alpha, w.p.p, w.q.p, x, w.r.p, w.part.p);
    rn = std::sqrt(fetch_scalar(c, nb));
    ++it;
    if (!std::isfinite(rn)) throw Error{MHX_ERR_NONFINITE, "CG: non-finite residual"};
  }
  // true residual
  spmv(c, A, x, w.q.p);
  hipLaunchKernelGGL(k_resid, g1, dim3(256), 0, c.stream, n, b, w.q.p, w.r.p);
  *relres = norm2(c, w.r.p, n) / bn;
  return it;
}

// ---- GMRES(m), right Jacobi preconditioning, CGS + DGKS re-orthogonalisation --
// h[k] = V_k . w for k < K  (V row-major: vector k at V + k*ld)
__global__ __launch_bounds__(256) void k_mdot(long n, int K, const double* __restrict__ V, long ld,
                                              const double* __restrict__ w, double* __restrict__ part) {
  // K <= 64 handled in chunks of 8 to bound registers
  for (int k0 = 0; k0 < K; k0 += 8) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int kk = min(8, K - k0);
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
      const double wi = w[i];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < kk) v[k] = fma(V[(long)(k0 + k) * ld + i], wi, v[k]);
    }
    __shared__ double sh[8][4];
    const int wv = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double s = wave_sum(v[k]);
      if (l == 0) sh[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x < kk) part[(long)(k0 + threadIdx.x) * kRedBlocks + blockIdx.x] =
        sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
    __syncthreads();
  }
}
// w -= sum_k h[k] V_k ; partial ||w||^2
__global__ __launch_bounds__(256) void k_mupdate(long n, int K, const double* __restrict__ V, long ld,
                                                 const double* __restrict__ h, double* __restrict__ w,
                                                 double* __restrict__ part) {
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    double wi = w[i];
    for (int k = 0; k < K; ++k) wi = fma(-h[k], V[(long)k * ld + i], wi);
    w[i] = wi;
    v[0] = fma(wi, wi, v[0]);
  }
  block_store<1>(v, part, kRedBlocks);
}
__global__ void k_scale_to(long n, double s, const double* __restrict__ x, double* __restrict__ y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = s * x[i];
}
// x += minv .* (V y)
__global__ void k_gm_update(long n, int K, const double* __restrict__ V, long ld, const double* __restrict__ y,
                            const double* __restrict__ minv, double* __restrict__ x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < K; ++k) s = fma(y[k], V[(long)k * ld + i], s);
  x[i] = fma(minv[i], s, x[i]);
}

static int gmres_solve(Ctx& c, const double* A, const double* b, double* x, int m, double tol, int maxit,
                       double* relres) {
  const long n = c.mesh.ndof;
  KrylovWork& w = c.kw;
  m = std::max(1, std::min(m, 64));
  w.r.alloc(n); w.w.alloc(n); w.minv.alloc(n); w.q.alloc(n);
  if (w.V.n < (size_t)(m + 1) * n) w.V.alloc((size_t)(m + 1) * n);
  ensure_red(c, m + 1);
  const dim3 g1((unsigned)((n + 255) / 256));
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_jacobi<double>, g1, dim3(256), 0, c.stream, n, c.rowptr.p, c.colidx.p, A, w.minv.p);
  const double bn = norm2(c, b, n);
  if (bn == 0.0) {
    MHX_HIP(hipMemsetAsync(x, 0, n * sizeof(double), c.stream));
    *relres = 0.0;
    return 0;
  }
  std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1), hcol(m + 1), y(m);
  int it = 0;
  double rel = 1.0;
  for (;;) {
    spmv(c, A, x, w.q.p);
    hipLaunchKernelGGL(k_resid, g1, dim3(256), 0, c.stream, n, b, w.q.p, w.r.p);
    const double beta = norm2(c, w.r.p, n);
    rel = beta / bn;
    if (!std::isfinite(rel)) throw Error{MHX_ERR_NONFINITE, "GMRES: non-finite residual"};
    if (rel <= tol || it >= maxit) break;
    hipLaunchKernelGGL(k_scale_to, g1, dim3(256), 0, c.stream, n, 1.0 / beta, w.r.p, w.V.p);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (; k < m && it < maxit;) {
      ScopedPhase ph(c.prof, "Krylov::gmres_iteration", c.stream);
      double* vk = w.V.p + (long)k * n;
      double* vk1 = w.V.p + (long)(k + 1) * n;
      spmv_t<double, double>(c, A, w.minv.p, vk, vk1);  // w = A M^-1 v_k, stored in V_{k+1}
      double wn0 = 0.0;
      // CGS pass 1
      std::fill(hcol.begin(), hcol.end(), 0.0);
      for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_mdot, dim3(nb), dim3(256), 0, c.stream, n, k + 1, w.V.p, n, vk1, w.part.p);
        hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, c.stream, w.part.p, nb, k + 1, 0, w.scal.p);
        std::vector<double> hh(k + 1);
        MHX_HIP(hipMemcpyAsync(hh.data(), w.scal.p, (k + 1) * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        MHX_HIP(hipStreamSynchronize(c.stream));
        if (pass == 0) {
          double s = 0.0;
          for (int j = 0; j <= k; ++j) s += hh[j] * hh[j];
          // ||w||^2 before projection = ||w_after||^2 + sum h^2 (Pythagoras), computed after update
          wn0 = s;
        }
        MHX_HIP(hipMemcpyAsync(w.scal.p + 128, hh.data(), (k + 1) * sizeof(double), hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(k_mupdate, dim3(nb), dim3(256), 0, c.stream, n, k + 1, w.V.p, n, w.scal.p + 128, vk1, w.part.p);
        const double wn2 = fetch_scalar(c, nb);
        for (int j = 0; j <= k; ++j) hcol[j] += hh[j];
        hcol[k + 1] = std::sqrt(wn2);
        // DGKS criterion: re-orthogonalise when the norm dropped by more than 1/sqrt(2)
        const double before = std::sqrt(wn2 + (pass == 0 ? wn0 : 0.0));
        if (pass == 0 && hcol[k + 1] > 0.7071067811865476 * before) break;
      }
      const double hn = hcol[k + 1];
      if (hn > 0.0)
        hipLaunchKernelGGL(k_scale_to, g1, dim3(256), 0, c.stream, n, 1.0 / hn, vk1, vk1);
      for (int j = 0; j <= k + 1; ++j) H[j * m + k] = hcol[j];
      for (int j = 0; j < k; ++j) {
        const double a0 = H[j * m + k], b0 = H[(j + 1) * m + k];
        H[j * m + k] = cs[j] * a0 + sn[j] * b0;
        H[(j + 1) * m + k] = -sn[j] * a0 + cs[j] * b0;
      }
      const double a0 = H[k * m + k], b0 = H[(k + 1) * m + k];
      const double den = std::sqrt(a0 * a0 + b0 * b0);
      cs[k] = den == 0.0 ? 1.0 : a0 / den;
      sn[k] = den == 0.0 ? 0.0 : b0 / den;
      H[k * m + k] = den;
      H[(k + 1) * m + k] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      ++k;
      ++it;
      if (std::fabs(g[k]) / bn <= 0.5 * tol || hn == 0.0) break;
    }
    for (int j = k - 1; j >= 0; --j) {
      double s = g[j];
      for (int l = j + 1; l < k; ++l) s -= H[j * m + l] * y[l];
      y[j] = s / H[j * m + j];
    }
    MHX_HIP(hipMemcpyAsync(w.scal.p, y.data(), k * sizeof(double), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_gm_update, g1, dim3(256), 0, c.stream, n, k, w.V.p, n, w.scal.p, w.minv.p, x);
    MHX_HIP(hipGetLastError());
  }
  *relres = rel;
  return it;
}

int krylov_solve(Ctx& c, const double* A, const double* b, double* x, const mhx_krylov_opts& o, mhx_solve_stats* st) {
  ScopedPhase ph(c.prof, "LinearAlgebraInterface::linearSolver", c.stream);
  double rel = 0.0;
  int it;
  const int maxit = o.max_iters > 0 ? o.max_iters : 100;
  const double tol = o.tol > 0 ? o.tol : 1e-7;
  if (o.solver == MHX_KRYLOV_CG) it = cg_solve(c, A, b, x, tol, maxit, &rel);
  else it = gmres_solve(c, A, b, x, o.restart > 0 ? o.restart : 50, tol, maxit, &rel);
  if (st) {
    st->iters += it;
    st->rel_res = rel;
  }
  if (!(rel <= tol)) return MHX_ERR_NOT_CONVERGED;
  return MHX_OK;
}

}  // namespace mhx
