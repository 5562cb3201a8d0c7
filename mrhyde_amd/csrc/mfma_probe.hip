// mfma_probe.hip — the dense B^T D B form of the Q2 hexahedron's element
// contraction on the FP64 matrix cores (v_mfma_f64_16x16x4f64), measured as
// the alternative to the sum-factorised column stage of the Q2 assembly
// kernels (k_thermal, k_q2box).
//
// Element matrix A = G diag(d) G^T: G is the 27 x 108 table of the Q2 basis
// at the 27 Gauss points (row i = node; column k = 4 q + c: dphi_i/dxi_c at
// point q for c < 3, phi_i(q) for c = 3) and d holds the element's 108 point
// coefficients (the thermal flux derivatives and weights of the sum-
// factorised kernels).  27 is padded to 32: 4 output tiles of 16 x 16, each
// the sum of 27 k-blocks of 4, i.e. 108 MFMA instructions (2048 FLOP each)
// per element, 78.7 k multiply-adds against the 11.7 k FMA of the
// sum-factorised form.  One wavefront per element at a time, G in LDS, the
// lane's 27 coefficients in registers.  Each element's matrix is reduced to
// the sum of its entries (sums[e]); the first n_full elements also store the
// whole padded matrix in the accumulator's order, which the host maps to
// rows and columns with the layout a one-instruction kernel reports and
// compares with G diag(d) G^T.

#include <cmath>
#include <vector>

#include "mhx_internal.hpp"

namespace mhx {

namespace {

// 1-D Q2 on [-1, 1], nodes -1, 0, 1; 3-point Gauss points -sqrt(3/5), 0, sqrt(3/5)
__host__ __device__ inline double q2_val(int a, double x) {
  return a == 0 ? 0.5 * x * (x - 1.0) : (a == 1 ? 1.0 - x * x : 0.5 * x * (x + 1.0));
}
__host__ __device__ inline double q2_der(int a, double x) { return a == 0 ? x - 0.5 : (a == 1 ? -2.0 * x : x + 0.5); }
__host__ __device__ inline double gauss3(int q) { return q == 0 ? -0.77459666924148337704 : (q == 1 ? 0.0 : 0.77459666924148337704); }

// G[i][4 q + c]
__host__ __device__ inline double q2_table(int i, int k) {
  const int q = k / 4, c = k % 4;
  const int in[3] = {i % 3, (i / 3) % 3, i / 9}, qn[3] = {q % 3, (q / 3) % 3, q / 9};
  double v = 1.0;
  for (int d = 0; d < 3; ++d) {
    const double x = gauss3(qn[d]);
    v *= (d == c) ? q2_der(in[d], x) : q2_val(in[d], x);
  }
  return v;
}

// deterministic coefficient of element e, slot k (a hash in [0.5, 1.5))
__host__ __device__ inline double probe_coef(long e, int k) {
  unsigned long long h = (unsigned long long)e * 0x9E3779B97F4A7C15ull + (unsigned long long)k * 0xBF58476D1CE4E5B9ull;
  h ^= h >> 31;
  h *= 0x94D049BB133111EBull;
  h ^= h >> 29;
  return 0.5 + (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void k_probe_fill(long n, double* __restrict__ d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n * 108) d[t] = probe_coef(t / 108, (int)(t % 108));
}

typedef double double4_t __attribute__((ext_vector_type(4)));

// Where the instruction puts D[i][j]: with A[i][0] = i, A[i][1] = 1 and
// B[0][j] = 1, B[1][j] = 100 j (k = 2, 3 zero), D[i][j] = i + 100 j, so
// out[4 l + v] names the row and column that lane l's slot v holds.
__global__ void k_mfma_layout(double* __restrict__ out) {
  const int l = threadIdx.x, r = l % 16, kk = l / 16;
  const double a = kk == 0 ? (double)r : (kk == 1 ? 1.0 : 0.0);
  const double b = kk == 0 ? 1.0 : (kk == 1 ? 100.0 * r : 0.0);
  double4_t acc = {0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
#pragma unroll
  for (int v = 0; v < 4; ++v) out[4 * l + v] = acc[v];
}

// The basis table G (rows padded 27 -> 32) in LDS, built by the block.
__device__ inline void fill_table(double (*sG)[108]) {
  for (int t = threadIdx.x; t < 32 * 108; t += blockDim.x) {
    const int i = t / 108, k = t % 108;
    sG[i][k] = i < 27 ? q2_table(i, k) : 0.0;
  }
  __syncthreads();
}

// Element matrix tile (ti, tj) of element e, accumulator lane l, slot v,
// stored in the accumulator's own order; the host maps it with the layout
// that k_mfma_layout reports.
__device__ inline void store_tile(double* full, long e, int tt, int l, const double4_t& acc) {
#pragma unroll
  for (int v = 0; v < 4; ++v) full[((e * 4 + tt) * 64 + l) * 4 + v] = acc[v];
}

// variant 0: G read from LDS for every instruction, one accumulator chain
// of 27 instructions per tile
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_mfma_lds(long n_elem, const double* __restrict__ d,
                                                  double* __restrict__ sums, long n_full,
                                                  double* __restrict__ full) {
  __shared__ double sG[32][108];
  fill_table(sG);
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64;
  const int r = l % 16, kk = l / 16;  // operand lane layout: A[r][kk], B[kk][r]
  for (long e = (long)blockIdx.x * 4 + wv; e < n_elem; e += (long)gridDim.x * 4) {
    double dk[27];
#pragma unroll
    for (int kb = 0; kb < 27; ++kb) dk[kb] = d[e * 108 + kb * 4 + kk];
    double s = 0.0;
#pragma unroll 1
    for (int tt = 0; tt < 4; ++tt) {
      const int ti = tt >> 1, tj = tt & 1;
      double4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kb = 0; kb < 27; ++kb) {
        const double a = sG[ti * 16 + r][kb * 4 + kk];
        const double b = sG[tj * 16 + r][kb * 4 + kk] * dk[kb];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
      s += (acc[0] + acc[1]) + (acc[2] + acc[3]);
      if (e < n_full) store_tile(full, e, tt, l, acc);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (l == 0) sums[e] = s;
  }
}

// variant 1: the lane's 54 entries of G held in registers for the whole
// kernel, and two accumulator chains (tiles (0, tj) and (1, tj), which share
// the B operand) interleaved
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_mfma_reg(long n_elem, const double* __restrict__ d,
                                                  double* __restrict__ sums, long n_full,
                                                  double* __restrict__ full) {
  __shared__ double sG[32][108];
  fill_table(sG);
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64;
  const int r = l % 16, kk = l / 16;
  double g0[27], g1[27];
#pragma unroll
  for (int kb = 0; kb < 27; ++kb) {
    g0[kb] = sG[r][kb * 4 + kk];
    g1[kb] = sG[16 + r][kb * 4 + kk];
  }
  for (long e = (long)blockIdx.x * 4 + wv; e < n_elem; e += (long)gridDim.x * 4) {
    double dk[27];
#pragma unroll
    for (int kb = 0; kb < 27; ++kb) dk[kb] = d[e * 108 + kb * 4 + kk];
    double s = 0.0;
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      double4_t c0 = {0.0, 0.0, 0.0, 0.0}, c1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kb = 0; kb < 27; ++kb) {
        const double b = (tj == 0 ? g0[kb] : g1[kb]) * dk[kb];
        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(g0[kb], b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(g1[kb], b, c1, 0, 0, 0);
      }
      s += ((c0[0] + c0[1]) + (c0[2] + c0[3])) + ((c1[0] + c1[1]) + (c1[2] + c1[3]));
      if (e < n_full) {
        store_tile(full, e, tj, l, c0);      // tile (0, tj): tt = tj
        store_tile(full, e, 2 + tj, l, c1);  // tile (1, tj): tt = 2 + tj
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (l == 0) sums[e] = s;
  }
}

}  // namespace

}  // namespace mhx

using namespace mhx;

extern "C" int mhx_mfma_probe(int64_t n_elem, int32_t reps, int32_t variant, double* ms_per_pass,
                              double* max_rel_err) {
  try {
    if (n_elem <= 0 || n_elem > (int64_t)1 << 24 || reps <= 0 || variant < 0 || variant > 1 || !ms_per_pass ||
        !max_rel_err)
      return MHX_ERR_ARG;
    const int64_t n_full = std::min<int64_t>(n_elem, 64);
    DevBuf<double> d, sums, full, lay;
    d.alloc((size_t)n_elem * 108);
    sums.alloc((size_t)n_elem);
    full.alloc((size_t)n_full * 4 * 64 * 4);
    lay.alloc(256);
    // stream and events released on every exit path (the MHX_HIP throws included)
    struct Res {
      hipStream_t s = nullptr;
      hipEvent_t a = nullptr, b = nullptr;
      ~Res() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        if (s) (void)hipStreamDestroy(s);
      }
    } res;
    MHX_HIP(hipStreamCreateWithFlags(&res.s, hipStreamNonBlocking));
    MHX_HIP(hipEventCreate(&res.a));
    MHX_HIP(hipEventCreate(&res.b));
    hipStream_t s = res.s;
    hipEvent_t a = res.a, b = res.b;
    hipLaunchKernelGGL(k_mfma_layout, dim3(1), dim3(64), 0, s, lay.p);
    std::vector<double> hl(256);
    MHX_HIP(hipMemcpyAsync(hl.data(), lay.p, 256 * sizeof(double), hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(k_probe_fill, dim3((unsigned)((n_elem * 108 + 255) / 256)), dim3(256), 0, s, (long)n_elem, d.p);
    const unsigned grid = (unsigned)std::min<int64_t>((n_elem + 3) / 4, 2048);
    auto kern = variant == 0 ? k_mfma_lds : k_mfma_reg;
    // checking pass (stores the first n_full matrices), which also warms up
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s, (long)n_elem, d.p, sums.p, (long)n_full, full.p);
    std::vector<double> hf((size_t)n_full * 4 * 64 * 4);
    MHX_HIP(hipMemcpyAsync(hf.data(), full.p, hf.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    MHX_HIP(hipEventRecord(a, s));
    for (int k = 0; k < reps; ++k)
      hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s, (long)n_elem, d.p, sums.p, 0L, full.p);
    MHX_HIP(hipEventRecord(b, s));
    MHX_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    MHX_HIP(hipEventElapsedTime(&ms, a, b));
    *ms_per_pass = ms / reps;
    std::vector<double> h(n_elem);
    MHX_HIP(hipMemcpyAsync(h.data(), sums.p, n_elem * sizeof(double), hipMemcpyDeviceToHost, s));
    MHX_HIP(hipStreamSynchronize(s));
    // the accumulator layout: every (row, column) of the 16 x 16 tile exactly once
    int row[256], colm[256], seen[256] = {0};
    for (int t = 0; t < 256; ++t) {
      const long code = std::lround(hl[t]);
      row[t] = (int)(code % 100);
      colm[t] = (int)(code / 100);
      if (std::fabs(hl[t] - (double)code) > 0 || row[t] > 15 || colm[t] > 15 || code < 0 || seen[row[t] * 16 + colm[t]]++)
        throw Error{MHX_ERR_DEVICE, "matrix-core accumulator layout is not a permutation of the tile"};
    }
    // the stored matrices against G diag(d) G^T (padding zero), and every
    // element's entry sum against sum_k d_k (sum_i G_ik)^2
    double G[27][108], col[108];
    for (int i = 0; i < 27; ++i)
      for (int k = 0; k < 108; ++k) G[i][k] = q2_table(i, k);
    for (int k = 0; k < 108; ++k) {
      col[k] = 0.0;
      for (int i = 0; i < 27; ++i) col[k] += G[i][k];
    }
    double err = 0.0;
    for (int64_t e = 0; e < n_full; ++e) {
      double scale = 0.0;
      for (int k = 0; k < 108; ++k)
        for (int i = 0; i < 27; ++i) scale = std::max(scale, probe_coef(e, k) * G[i][k] * G[i][k]);
      for (int tt = 0; tt < 4; ++tt)
        for (int t = 0; t < 256; ++t) {
          const int i = (tt >> 1) * 16 + row[t], j = (tt & 1) * 16 + colm[t];
          double ref = 0.0;
          if (i < 27 && j < 27)
            for (int k = 0; k < 108; ++k) ref += G[i][k] * probe_coef(e, k) * G[j][k];
          err = std::max(err, std::fabs(hf[(e * 4 + tt) * 256 + t] - ref) / scale);
        }
    }
    for (int64_t e = 0; e < n_elem; ++e) {
      double ref = 0.0;
      for (int k = 0; k < 108; ++k) ref += probe_coef(e, k) * col[k] * col[k];
      err = std::max(err, std::fabs(h[e] - ref) / std::max(std::fabs(ref), 1e-300));
    }
    *max_rel_err = err;
    return MHX_OK;
  } catch (const Error& e) {
    return e.code;
  }
}
