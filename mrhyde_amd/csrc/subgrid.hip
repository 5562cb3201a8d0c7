// subgrid.hip — batched multiscale subgrid solves: thousands of independent
// small Q1 thermal problems assembled and solved (Jacobi-PCG to a relative
// residual) in ONE persistent launch.
//
// Replaces the per-macro-group loop of the reference multiscale layer, which
// solves one subgrid system at a time with its own assembly and Krylov solve:
//   MultiscaleManager::evaluateMacroMicroMacroMap   multiscaleManager.cpp:866-1188
//   SubGridDtN::subgridSolver                        subgridDtN.cpp:985-1121
//   SubGridDtN_Solver::nonlinearSolver / assembleJacobianResidual
//                                                    subgridDtN_solver.cpp:903-1043, :681-902
// with the PCG of SolverManager::PCG (solverManager_solvers.hpp:812-905).
//
// Instance: a box [0,Lx]x[0,Ly]x[0,Lz] (the unit cube by default) cut into
// n^3 Q1 hexahedra (n per instance), element conductivity k_e, optional
// element source f_e, Dirichlet data g on every boundary node.  The system is
// J du = -F(u0) with u0 = g on the boundary (reference strong-DBC convention)
// and F(u) = A u - b, b_i = sum over the elements e of node i of f_e |e| / 8,
// solved on the interior nodes; the result is u = u0 + du.
//
// Mapping: one workgroup owns one instance at a time and pulls the next from
// a global counter (largest instances first) until the queue is empty, so
// every workgroup reaches the exit.  The instance lives on chip: the search
// direction p and the element conductivities in LDS, each thread's share of
// x, r and the Jacobi scaling in registers.  The operator is applied node by
// node from the 27 neighbouring p values and the 8 surrounding element
// conductivities with the closed-form Q1 stiffness of the instance's box
// element (exact for the 2x2x2 Gauss rule the reference uses), i.e. the
// assembled 27-point row is formed on the fly instead of being stored.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mhx_internal.hpp"

namespace mhx {

// ---- box trilinear stiffness -------------------------------------------------
// On a box element of sides (hx, hy, hz) the Q1 stiffness is
//   K = cx S(x)M(x)M + cy M(x)S(x)M + cz M(x)M(x)S,
//   cx = hy hz / hx, cy = hx hz / hy, cz = hx hy / hz,
// with the 1-D S = [1 -1; -1 1] and M = [1/3 1/6; 1/6 1/3].  An entry depends
// only on which of the two nodes' coordinates differ: kv[dx + 2 dy + 4 dz].
inline void box_stencil(double hx, double hy, double hz, double kv[8]) {
  const double cx = hy * hz / hx, cy = hx * hz / hy, cz = hx * hy / hz;
  const double S[2] = {1.0, -1.0}, M[2] = {1.0 / 3.0, 1.0 / 6.0};
  for (int p = 0; p < 8; ++p) {
    const int dx = p & 1, dy = (p >> 1) & 1, dz = (p >> 2) & 1;
    kv[p] = cx * S[dx] * M[dy] * M[dz] + cy * M[dx] * S[dy] * M[dz] + cz * M[dx] * M[dy] * S[dz];
  }
}

// the unit-cube table (hx = hy = hz = 1): K_h = h K_unit for a cube of side h
struct K8 {
  double v[8][8];
};
constexpr K8 make_k8() {
  K8 k{};
  const double M[2][2] = {{1.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 1.0 / 3.0}};
  const double S[2][2] = {{1.0, -1.0}, {-1.0, 1.0}};
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) {
      const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
      const int bx = b & 1, by = (b >> 1) & 1, bz = (b >> 2) & 1;
      k.v[a][b] = S[ax][bx] * M[ay][by] * M[az][bz] + M[ax][bx] * S[ay][by] * M[az][bz] +
                  M[ax][bx] * M[ay][by] * S[az][bz];
    }
  return k;
}
constexpr K8 kK8 = make_k8();

constexpr int kSgMaxN = 16;                                   // elements per axis
constexpr int kSgMaxNodes = (kSgMaxN + 1) * (kSgMaxN + 1) * (kSgMaxN + 1);
constexpr int kSgMaxElems = kSgMaxN * kSgMaxN * kSgMaxN;

struct SgInst {
  int n;       // elements per axis
  int k_off;   // offset of the instance's element coefficients
  int x_off;   // offset of the instance's nodal arrays
  int pad;
  double kv[8];  // element stiffness entries by coordinate-difference pattern
  double vol8;   // element volume / 8: the share of a node in a constant source
};

struct SgArgs {
  const SgInst* inst;
  const int* order;      // queue position -> instance (largest first)
  int n_inst;
  const double* k;
  const double* src;     // element sources (nullptr: none)
  const double* g;
  double* x;
  int* iters;
  double* relres;
  int* refinements;
  unsigned int* queue;
  double tol;
  int max_iters;
  int fp32;              // FP32 inner CG inside FP64 iterative refinement
  double inner_tol;
  int max_refine;
};

__global__ void k_sg_reset(unsigned int* q) { *q = 0u; }

// q = A p at node (ix,iy,iz) (interior), p and k in LDS.  BOX (the general
// instance: a box, possibly with sources): kv holds the box element's
// stiffness entries by coordinate-difference pattern; else the instance is
// the unit cube without sources and its element stiffness h K_unit (h = 1/n)
// comes from the compile-time table.
template <class T, bool BOX>
__device__ __forceinline__ T sg_apply(const T* __restrict__ p, const T* __restrict__ kk, int n, const T (&kv)[8],
                                      int ix, int iy, int iz) {
  const int n1 = n + 1;
  T P[3][3][3];
#pragma unroll
  for (int dz = 0; dz < 3; ++dz)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) P[dz][dy][dx] = p[(ix - 1 + dx) + n1 * ((iy - 1 + dy) + n1 * (iz - 1 + dz))];
  T acc = 0;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
    const T ke = kk[(ix - 1 + ox) + n * ((iy - 1 + oy) + n * (iz - 1 + oz))];
    const int a = (1 - ox) + 2 * (1 - oy) + 4 * (1 - oz);
    T s = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int bx = b & 1, by = (b >> 1) & 1, bz = (b >> 2) & 1;
      const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
      const T kab = BOX ? kv[(ax ^ bx) | ((ay ^ by) << 1) | ((az ^ bz) << 2)] : (T)kK8.v[a][b];
      s += kab * P[oz + bz][oy + by][ox + bx];
    }
    acc += ke * s;
  }
  if constexpr (BOX) return acc;
  else return acc * ((T)1 / (T)n);
}

// b_i / vol8: the sum of the element sources around node (ix,iy,iz)
__device__ __forceinline__ double sg_src(const double* __restrict__ f, int n, int ix, int iy, int iz) {
  double s = 0.0;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
    s += f[(ix - 1 + ox) + n * ((iy - 1 + oy) + n * (iz - 1 + oz))];
  }
  return s;
}

template <class T, bool BOX>
__device__ __forceinline__ T sg_diag(const T* __restrict__ kk, int n, T kv0, int ix, int iy, int iz) {
  T d = 0;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
    const int a = (1 - ox) + 2 * (1 - oy) + 4 * (1 - oz);
    const T ke = kk[(ix - 1 + ox) + n * ((iy - 1 + oy) + n * (iz - 1 + oz))];
    d += BOX ? ke : ke * (T)kK8.v[a][a];
  }
  if constexpr (BOX) return d * kv0;  // once per node at set-up
  else return d / (T)n;
}

// Sum of two values over the workgroup, returned to every thread.
template <int WG, class T>
__device__ __forceinline__ void wg_sum2(T& a, T& b, T* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  a = 0;
  b = 0;
#pragma unroll
  for (int i = 0; i < WG / 64; ++i) {
    a += red[2 * i];
    b += red[2 * i + 1];
  }
  __syncthreads();
}

// Jacobi-PCG on the interior nodes (SolverManager::PCG algorithm):
//   x += CG solution of A x = r_in (r_in given in the thread's slots),
// stopping when ||r|| / ||r_in|| <= tol.  Returns iterations; rr_out = ||r||^2.
template <int WG, int S, class T, bool BOX>
__device__ __forceinline__ int sg_pcg(int n, const T* kk, const T (&kv)[8], T* p_lds, T* red, T (&x)[S], T (&r)[S],
                                      const T (&dinv)[S], bool (&in)[S], int (&node)[S], double tol, int max_iters,
                                      T& rr_out) {
  const int n1 = n + 1, nn = n1 * n1 * n1;
  T pp[S];
  T rr = 0, rz = 0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    rr += r[s] * r[s];
    rz += r[s] * (dinv[s] * r[s]);  // z = D^-1 r (dinv is 0 off the interior)
  }
  wg_sum2<WG>(rr, rz, red);
  const T rr0 = rr;
  int it = 0;
  if (rr0 > (T)0) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      pp[s] = dinv[s] * r[s];
      if (node[s] < nn) p_lds[node[s]] = pp[s];
    }
    __syncthreads();
    const T tol2 = (T)(tol * tol);
    while (it < max_iters) {
      // The node indices are made opaque once per iteration, so the 27-point
      // gather addresses are recomputed inside the loop instead of being
      // hoisted for every slot and held across iterations.
#pragma unroll
      for (int s = 0; s < S; ++s) __asm__ volatile("" : "+v"(node[s]));
      T q[S];
      T pq = 0, dummy = 0;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        // one slot's 27-point gather at a time (bounds the live registers)
        __builtin_amdgcn_sched_barrier(0);
        q[s] = 0;
        if (in[s]) {
          const int v = node[s];
          q[s] = sg_apply<T, BOX>(p_lds, kk, n, kv, v % n1, (v / n1) % n1, v / (n1 * n1));
        }
        pq += pp[s] * q[s];
      }
      wg_sum2<WG>(pq, dummy, red);
      const T alpha = rz / pq;
      T rr1 = 0, rz1 = 0;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        x[s] += alpha * pp[s];
        r[s] -= alpha * q[s];
        rr1 += r[s] * r[s];
        rz1 += r[s] * (dinv[s] * r[s]);
      }
      wg_sum2<WG>(rr1, rz1, red);
      ++it;
      rr = rr1;
      if (!(rr > tol2 * rr0)) break;  // also leaves on NaN
      const T beta = rz1 / rz;
      rz = rz1;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        pp[s] = dinv[s] * r[s] + beta * pp[s];
        if (node[s] < nn) p_lds[node[s]] = pp[s];
      }
      __syncthreads();
    }
  }
  rr_out = rr;
  return it;
}

template <int WG, int S, bool MIXED, bool BOX>
__global__ __launch_bounds__(WG) void k_subgrid(SgArgs a) {
  __shared__ double p64[kSgMaxNodes];
  __shared__ double k64[kSgMaxElems];
  __shared__ double red[2 * (WG / 64)];
  __shared__ int s_next;
  float* p32 = reinterpret_cast<float*>(p64);       // FP32 views (first half of each array)
  float* k32 = reinterpret_cast<float*>(p64) + kSgMaxNodes;  // second half of p64 storage
  __shared__ float red32[2 * (WG / 64)];

  for (;;) {
    if (threadIdx.x == 0) s_next = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    // (wave-uniform: the instance's description is then read with scalar
    // loads and its stencil entries stay in scalar registers)
    const int qpos = s_next;
    __syncthreads();
    if (qpos >= a.n_inst) break;
    const int id = a.order[qpos];
    const SgInst in_ = a.inst[id];
    const int n = in_.n, n1 = n + 1, nn = n1 * n1 * n1, ne = n * n * n;
    const double* kg = a.k + in_.k_off;
    const double* fg = a.src ? a.src + in_.k_off : nullptr;
    const double* gg = a.g + in_.x_off;
    double kv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) kv[i] = in_.kv[i];

    for (int e = threadIdx.x; e < ne; e += WG) k64[e] = kg[e];
    int node[S];
    bool in[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int v = threadIdx.x + s * WG;
      node[s] = v < nn ? v : nn;
      const int ix = v % n1, iy = (v / n1) % n1, iz = v / (n1 * n1);
      const bool bnd = ix == 0 || iy == 0 || iz == 0 || ix == n || iy == n || iz == n;
      in[s] = v < nn && !bnd;
      if (v < nn) p64[v] = bnd ? gg[v] : 0.0;  // u0: the Dirichlet lift
    }
    __syncthreads();
    // b - A u0 = -F(u0) on interior rows (assembled residual with the Dirichlet lift)
    double r[S], x[S], dinv[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      __builtin_amdgcn_sched_barrier(0);
      x[s] = 0.0;
      r[s] = 0.0;
      dinv[s] = 0.0;
      if (in[s]) {
        const int v = node[s];
        const int ix = v % n1, iy = (v / n1) % n1, iz = v / (n1 * n1);
        const double b = (BOX && fg) ? in_.vol8 * sg_src(fg, n, ix, iy, iz) : 0.0;
        r[s] = b - sg_apply<double, BOX>(p64, k64, n, kv, ix, iy, iz);
        dinv[s] = 1.0 / sg_diag<double, BOX>(k64, n, kv[0], ix, iy, iz);
      }
    }
    __syncthreads();
    int its = 0, refs = 0;
    double rr = 0.0, bb = 0.0, dummy = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) bb += r[s] * r[s];
    wg_sum2<WG>(bb, dummy, red);
    if constexpr (!MIXED) {
      its = sg_pcg<WG, S, double, BOX>(n, k64, kv, p64, red, x, r, dinv, in, node, a.tol, a.max_iters, rr);
    } else {
      // FP32 inner PCG inside FP64 iterative refinement: the residual, the
      // accumulated correction and the convergence test stay FP64.
      const double target2 = a.tol * a.tol * bb;
      rr = bb;
      while (refs < a.max_refine && rr > target2 && its < a.max_iters) {
        const double scale = std::sqrt(rr);
        // FP32 views: p32 in the lower half of p64's storage, k32 in its upper half
        __syncthreads();
        for (int e = threadIdx.x; e < ne; e += WG) k32[e] = (float)k64[e];
        float xf[S], rf[S], df[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
          xf[s] = 0.f;
          rf[s] = (float)(r[s] / scale);
          df[s] = (float)dinv[s];
        }
        __syncthreads();
        float rrf = 0.f, kvf[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) kvf[i] = (float)kv[i];
        its += sg_pcg<WG, S, float, BOX>(n, k32, kvf, p32, red32, xf, rf, df, in, node, a.inner_tol, a.max_iters - its, rrf);
        ++refs;
        // x += scale * xf ; r = b - A (u0 + x) in FP64
        __syncthreads();
#pragma unroll
        for (int s = 0; s < S; ++s) {
          x[s] += scale * (double)xf[s];
          const int v = node[s];
          if (v < nn) p64[v] = (in[s] ? 0.0 : gg[v]) + x[s];
        }
        __syncthreads();
        double rr2 = 0.0, d2 = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          __builtin_amdgcn_sched_barrier(0);
          if (in[s]) {
            const int v = node[s];
            const int ix = v % n1, iy = (v / n1) % n1, iz = v / (n1 * n1);
            const double b = (BOX && fg) ? in_.vol8 * sg_src(fg, n, ix, iy, iz) : 0.0;
            r[s] = b - sg_apply<double, BOX>(p64, k64, n, kv, ix, iy, iz);
          }
          rr2 += r[s] * r[s];
        }
        wg_sum2<WG>(rr2, d2, red);
        rr = rr2;
      }
    }
    // u = u0 + du
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int v = node[s];
      if (v < nn) a.x[in_.x_off + v] = (in[s] ? 0.0 : gg[v]) + x[s];
    }
    if (threadIdx.x == 0) {
      a.iters[id] = its;
      a.relres[id] = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
      if (a.refinements) a.refinements[id] = refs;
    }
    __syncthreads();
  }
}

}  // namespace mhx

// =============================================================================
// host side: batch object, autotuner, graph capture, C ABI
// =============================================================================
namespace mhx {

struct Batch {
  int n_inst = 0;
  long total_nodes = 0, total_elems = 0;
  std::vector<SgInst> h_inst;
  std::vector<int> h_order;
  DevBuf<SgInst> inst;
  DevBuf<int> order, iters, refs;
  DevBuf<double> k, src, g, x, relres;
  DevBuf<unsigned int> queue;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_wg = 0;
  bool box = false;  // instances with their own side lengths or sources (else unit cubes, no source)
  int wg = 0;
  int grid[2] = {0, 0};  // per WG variant (512, 1024)
  mhx_batch_desc desc{};
  double last_ms = 0.0;
  std::string err;
  ~Batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static SgArgs make_args(Batch& b) {
  SgArgs a{};
  a.inst = b.inst.p;
  a.order = b.order.p;
  a.n_inst = b.n_inst;
  a.k = b.k.p;
  a.src = b.src.p;
  a.g = b.g.p;
  a.x = b.x.p;
  a.iters = b.iters.p;
  a.relres = b.relres.p;
  a.refinements = b.refs.p;
  a.queue = b.queue.p;
  a.tol = b.desc.tol;
  a.max_iters = b.desc.max_iters;
  a.fp32 = b.desc.precision == MHX_BATCH_FP32_REFINED;
  a.inner_tol = b.desc.inner_tol;
  a.max_refine = b.desc.max_refine;
  return a;
}

static void enqueue(Batch& b, int wg) {
  SgArgs a = make_args(b);
  hipLaunchKernelGGL(k_sg_reset, dim3(1), dim3(1), 0, b.stream, b.queue.p);
  constexpr int S1024 = (kSgMaxNodes + 1023) / 1024, S512 = (kSgMaxNodes + 511) / 512;
  auto go = [&](auto BX) {
    constexpr bool X = decltype(BX)::value;
    if (wg == 1024) {
      if (a.fp32)
        hipLaunchKernelGGL((k_subgrid<1024, S1024, true, X>), dim3(b.grid[1]), dim3(1024), 0, b.stream, a);
      else
        hipLaunchKernelGGL((k_subgrid<1024, S1024, false, X>), dim3(b.grid[1]), dim3(1024), 0, b.stream, a);
    } else {
      if (a.fp32)
        hipLaunchKernelGGL((k_subgrid<512, S512, true, X>), dim3(b.grid[0]), dim3(512), 0, b.stream, a);
      else
        hipLaunchKernelGGL((k_subgrid<512, S512, false, X>), dim3(b.grid[0]), dim3(512), 0, b.stream, a);
    }
  };
  if (b.box) go(std::true_type{});
  else go(std::false_type{});
  MHX_HIP(hipGetLastError());
}

// a pair of timing events destroyed on every exit path, including the
// MHX_HIP throws between creation and the end of the timed region
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  EventPair() {
    MHX_HIP(hipEventCreate(&a));
    hipError_t e = hipEventCreate(&b);
    if (e != hipSuccess) {
      (void)hipEventDestroy(a);
      throw Error{MHX_ERR_DEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e)};
    }
  }
  EventPair(const EventPair&) = delete;
  EventPair& operator=(const EventPair&) = delete;
  ~EventPair() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  float elapsed_ms() const {
    float ms = 0.f;
    MHX_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
  }
};

static double timed_run(Batch& b, int wg) {
  EventPair ev;
  MHX_HIP(hipEventRecord(ev.a, b.stream));
  enqueue(b, wg);
  MHX_HIP(hipEventRecord(ev.b, b.stream));
  MHX_HIP(hipEventSynchronize(ev.b));
  return ev.elapsed_ms();
}

static void setup_grid(Batch& b) {
  int dev = 0, cus = 0;
  MHX_HIP(hipGetDevice(&dev));
  MHX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int occ512 = 0, occ1024 = 0;
  MHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &occ512, (const void*)k_subgrid<512, (kSgMaxNodes + 511) / 512, false, false>, 512, 0));
  MHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &occ1024, (const void*)k_subgrid<1024, (kSgMaxNodes + 1023) / 1024, false, false>, 1024, 0));
  MHX_REQUIRE(occ512 > 0 && occ1024 > 0, MHX_ERR_DEVICE, "subgrid kernel does not fit on a CU");
  b.grid[0] = std::max(1, std::min(b.n_inst, cus * occ512));
  b.grid[1] = std::max(1, std::min(b.n_inst, cus * occ1024));
}

}  // namespace mhx

using namespace mhx;

struct mhx_batch : public Batch {};

extern "C" {

int mhx_batch_create(const mhx_batch_desc* d, mhx_batch** out) {
  mhx_batch* b = nullptr;
  try {
    MHX_REQUIRE(d && out && d->n_inst > 0 && d->n_axis && d->k && d->g, MHX_ERR_ARG, "bad batch description");
    *out = nullptr;
    b = new mhx_batch();
    b->desc = *d;
    if (b->desc.tol <= 0) b->desc.tol = 1e-10;
    if (b->desc.max_iters <= 0) b->desc.max_iters = 10000;
    if (b->desc.inner_tol <= 0) b->desc.inner_tol = 1e-5;
    if (b->desc.max_refine <= 0) b->desc.max_refine = 20;
    MHX_REQUIRE(b->desc.precision == MHX_BATCH_FP64 || b->desc.precision == MHX_BATCH_FP32_REFINED, MHX_ERR_ARG,
                "unknown batch precision");
    b->n_inst = d->n_inst;
    b->box = d->extent != nullptr || d->src != nullptr;
    b->h_inst.resize(b->n_inst);
    long ko = 0, xo = 0;
    for (int i = 0; i < b->n_inst; ++i) {
      const int n = d->n_axis[i];
      MHX_REQUIRE(n >= 2 && n <= kSgMaxN, MHX_ERR_UNSUPPORTED, "subgrid instances need 2..16 elements per axis");
      double L[3] = {1.0, 1.0, 1.0};
      if (d->extent)
        for (int c = 0; c < 3; ++c) {
          L[c] = d->extent[3 * (long)i + c];
          MHX_REQUIRE(std::isfinite(L[c]) && L[c] > 0.0, MHX_ERR_ARG, "subgrid box extents must be positive");
        }
      SgInst& si = b->h_inst[i];
      si = SgInst{};
      si.n = n, si.k_off = (int)ko, si.x_off = (int)xo;
      const double hx = L[0] / n, hy = L[1] / n, hz = L[2] / n;
      box_stencil(hx, hy, hz, si.kv);
      si.vol8 = hx * hy * hz / 8.0;
      ko += (long)n * n * n;
      xo += (long)(n + 1) * (n + 1) * (n + 1);
    }
    MHX_REQUIRE(xo < (1L << 31), MHX_ERR_UNSUPPORTED, "batch too large");
    b->total_elems = ko;
    b->total_nodes = xo;
    b->h_order.resize(b->n_inst);
    for (int i = 0; i < b->n_inst; ++i) b->h_order[i] = i;
    std::stable_sort(b->h_order.begin(), b->h_order.end(),
                     [&](int p, int q) { return b->h_inst[p].n > b->h_inst[q].n; });
    MHX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->inst.alloc(b->n_inst);
    b->order.alloc(b->n_inst);
    b->iters.alloc(b->n_inst);
    b->refs.alloc(b->n_inst);
    b->relres.alloc(b->n_inst);
    b->k.alloc(ko);
    if (d->src) b->src.alloc(ko);
    b->g.alloc(xo);
    b->x.alloc(xo);
    b->queue.alloc(1);
    MHX_HIP(hipMemcpyAsync(b->inst.p, b->h_inst.data(), b->n_inst * sizeof(SgInst), hipMemcpyHostToDevice, b->stream));
    MHX_HIP(hipMemcpyAsync(b->order.p, b->h_order.data(), b->n_inst * sizeof(int), hipMemcpyHostToDevice, b->stream));
    MHX_HIP(hipMemcpyAsync(b->k.p, d->k, ko * sizeof(double), hipMemcpyHostToDevice, b->stream));
    if (d->src) MHX_HIP(hipMemcpyAsync(b->src.p, d->src, ko * sizeof(double), hipMemcpyHostToDevice, b->stream));
    MHX_HIP(hipMemcpyAsync(b->g.p, d->g, xo * sizeof(double), hipMemcpyHostToDevice, b->stream));
    MHX_HIP(hipStreamSynchronize(b->stream));
    setup_grid(*b);
    b->wg = d->wg_size == 512 || d->wg_size == 1024 ? d->wg_size : 0;
    if (!b->wg) {
      // autotuner: time each workgroup shape on this batch (after one warm-up) and keep the fastest
      double best = 1e300;
      for (int wg : {512, 1024}) {
        (void)timed_run(*b, wg);
        const double t = timed_run(*b, wg);
        if (t < best) {
          best = t;
          b->wg = wg;
        }
      }
    }
    *out = b;
    return MHX_OK;
  } catch (const Error& e) {
    delete b;
    return e.code;
  }
}

void mhx_batch_destroy(mhx_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  delete b;
}

int mhx_batch_solve(mhx_batch* b, int32_t flags) {
  (void)flags;
  try {
    if (!b) return MHX_ERR_ARG;
    EventPair ev;
    if (b->desc.use_graph) {
      if (!b->exec || b->graph_wg != b->wg) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        MHX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        enqueue(*b, b->wg);
        MHX_HIP(hipStreamEndCapture(b->stream, &b->graph));
        MHX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_wg = b->wg;
      }
      MHX_HIP(hipEventRecord(ev.a, b->stream));
      MHX_HIP(hipGraphLaunch(b->exec, b->stream));
    } else {
      MHX_HIP(hipEventRecord(ev.a, b->stream));
      enqueue(*b, b->wg);
    }
    MHX_HIP(hipEventRecord(ev.b, b->stream));
    MHX_HIP(hipEventSynchronize(ev.b));
    b->last_ms = ev.elapsed_ms();
    return MHX_OK;
  } catch (const Error& e) {
    if (b) b->err = e.msg;
    return e.code;
  }
}

int mhx_batch_results(mhx_batch* b, double* x, int32_t* iters, double* relres, int32_t* refinements) {
  try {
    if (!b) return MHX_ERR_ARG;
    if (x) MHX_HIP(hipMemcpyAsync(x, b->x.p, b->total_nodes * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (iters) MHX_HIP(hipMemcpyAsync(iters, b->iters.p, b->n_inst * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    if (relres) MHX_HIP(hipMemcpyAsync(relres, b->relres.p, b->n_inst * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (refinements)
      MHX_HIP(hipMemcpyAsync(refinements, b->refs.p, b->n_inst * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    MHX_HIP(hipStreamSynchronize(b->stream));
    return MHX_OK;
  } catch (const Error& e) {
    b->err = e.msg;
    return e.code;
  }
}

int mhx_batch_info(const mhx_batch* b, int64_t* total_nodes, int32_t* wg_size, double* last_ms) {
  if (!b) return MHX_ERR_ARG;
  if (total_nodes) *total_nodes = b->total_nodes;
  if (wg_size) *wg_size = b->wg;
  if (last_ms) *last_ms = b->last_ms;
  return MHX_OK;
}

int mhx_batch_set_wg(mhx_batch* b, int32_t wg) {
  if (!b || (wg != 512 && wg != 1024)) return MHX_ERR_ARG;
  b->wg = wg;
  return MHX_OK;
}

}  // extern "C"
