// amg.hip — smoothed-aggregation algebraic multigrid on the node lattice
// (CDNA4, gfx950): hierarchy set-up on the device and the V-cycle used as a
// Krylov preconditioner.
//
// Replaces LinearAlgebraInterface::buildAMGPreconditioner
// (linearAlgebraInterface_solvers.hpp:201-253), which hands the Jacobian to
// MueLu::CreateTpetraPreconditioner with "multigrid algorithm" sa,
// "aggregation: type" uncoupled (drop scheme classical, nothing dropped),
// "smoother: type" CHEBYSHEV of degree 2 with eigenvalue ratio 7 and a zero
// starting solution, "coarse: max size" 500, no repartitioning, no reuse.
// MueLu (Trilinos) is not in the reference tree.  What is kept is the
// published smoothed-aggregation method (Vanek, Mandel, Brezina, Computing
// 56 (1996)) with MueLu's defaults for what the reference leaves unset:
//   - tentative prolongator T: the near-null space (one constant vector per
//     field, zero on strong-Dirichlet rows) restricted to each aggregate and
//     orthonormalised, which for one column per field is a normalisation;
//     the coarse near-null space is the column norm;
//   - P = (I - omega D^-1 A) T with omega = (4/3) / lambda, lambda the
//     10-step power-iteration estimate of rho(D^-1 A);
//   - R = P^T and A_c = R A P;
//   - Chebyshev smoothing on D^-1 A over [lmax / 7, lmax], lmax = 1.1 lambda,
//     before and after the coarse correction; V-cycle; the coarsest level
//     (at most 500 rows, or after 10 levels) is solved with its dense inverse.
// The aggregation differs from MueLu's greedy one in form, not in kind: on a
// lattice, uncoupled aggregation builds each aggregate from a root and its
// graph neighbours, i.e. (2r+1)^d node blocks for a stencil of reach r.  Here
// the aggregates ARE those lattice blocks (the last block along an axis takes
// the remainder), so every pattern of the hierarchy is a separable box known
// in closed form, and the Galerkin products need no hashing or sorting: each
// row accumulates into a fixed slot box in LDS, in a fixed order, so the
// hierarchy is bit-reproducible.
#include <algorithm>
#include <climits>
#include <cmath>
#include <memory>
#include <type_traits>

#include "lattice.hpp"
#include "mhx_internal.hpp"

namespace mhx {

constexpr int kAmgMaxLevels = 10;      // MueLu "max levels"
// MueLu's "coarse: max size" is 500 with a sparse direct coarse solver; the
// coarsest solve here is a dense inverse built by one workgroup in O(n^3), so
// the hierarchy coarsens on until at most 128 rows remain (one level more at
// most, since a lattice level shrinks 27x or more per step in 3-D)
constexpr long kAmgCoarseMax = 128;
constexpr long kAmgDenseMax = 512;     // Gauss-Jordan kernel limit (2n columns <= 1024 threads)
constexpr int kAmgPowerIters = 10;     // "sa: eigenvalue estimate num iterations"
constexpr double kAmgDamping = 4.0 / 3.0;
constexpr double kChebRatio = 7.0;     // "chebyshev: ratio eigenvalue"
constexpr double kChebBoost = 1.1;     // Ifpack2 Chebyshev boost of the estimate
constexpr int kAmgMaxFields = 4;
// smoothers: Chebyshev(2) on D^-1 A (the reference's MueLu setting), or
// damped sweeps of nodal block Jacobi (the nf x nf block of each node
// inverted), for coupled multi-field systems whose point diagonal does not
// bound the spectrum (stabilised Navier-Stokes).  On the developed 48^3 and
// 64^3 cavity flows damping above 0.5 makes the cycle diverge; three sweeps
// at 0.4 is the default.
constexpr int kSmootherChebyshev = 0, kSmootherBlockJacobi = 1, kSmootherBlockGaussSeidel = 2;
constexpr double kBjOmega = 0.4;  // defaults; options amg.sweeps and amg.omega_pct
constexpr int kBjSweeps = 3;
constexpr int kAmgBlk = 128;           // threads per workgroup of the set-up kernels

// ---------------------------------------------------------------------------
// lattice axis tables of one level.  Node p of an axis of n nodes lies in
// aggregate agg(p) = min(p / b, na - 1), b = 2r + 1, na = max(1, n / b).  The
// prolongator row of p spans aggregates plo(p) .. plo(p) + pext(p) - 1 (at most
// 2), the row of A P spans aplo(p) .. + apext(p) - 1 (at most 3), the
// restriction row of aggregate I spans fine nodes rlo(I) .. + rext(I) - 1, and
// the coarse operator row of I spans clo(I) .. + cext(I) - 1 (reach rc).  The
// s* arrays are prefix sums of the extents (closed-form row pointers).
// ---------------------------------------------------------------------------
enum AxisTable { kAgg, kPlo, kPext, kSp, kAplo, kApext, kRlo, kRext, kSr, kClo, kCext, kSc, kNumTables };

struct AxisDev {
  const int32_t* t[kNumTables];
  int n, na, b;
};
struct Geo {
  AxisDev ax[3];
  int nf;
};

struct AxisHost {
  int n = 1, na = 1, b = 1;
  std::vector<int32_t> t[kNumTables];
};

static void make_axis(AxisHost& a, int n, int r, int rc) {
  a.n = n;
  a.b = 2 * r + 1;
  a.na = std::max(1, n / a.b);
  const int b = a.b, na = a.na;
  for (auto& v : a.t) v.clear();
  auto& agg = a.t[kAgg];
  agg.resize(n);
  for (int p = 0; p < n; ++p) agg[p] = std::min(p / b, na - 1);
  a.t[kPlo].resize(n); a.t[kPext].resize(n); a.t[kAplo].resize(n); a.t[kApext].resize(n);
  a.t[kSp].assign(n + 1, 0);
  for (int p = 0; p < n; ++p) {
    const int lo = agg[std::max(0, p - r)];
    a.t[kPlo][p] = lo;
    a.t[kPext][p] = std::min(lo + 1, na - 1) - lo + 1;
    const int alo = agg[std::max(0, p - 2 * r)];
    a.t[kAplo][p] = alo;
    a.t[kApext][p] = std::min(alo + 2, na - 1) - alo + 1;
    a.t[kSp][p + 1] = a.t[kSp][p] + a.t[kPext][p];
  }
  std::vector<int> rlo(na, INT_MAX), rhi(na, -1);
  for (int p = 0; p < n; ++p)
    for (int I = a.t[kPlo][p]; I < a.t[kPlo][p] + a.t[kPext][p]; ++I) {
      rlo[I] = std::min(rlo[I], p);
      rhi[I] = std::max(rhi[I], p);
    }
  a.t[kRlo].resize(na); a.t[kRext].resize(na); a.t[kClo].resize(na); a.t[kCext].resize(na);
  a.t[kSr].assign(na + 1, 0);
  a.t[kSc].assign(na + 1, 0);
  for (int I = 0; I < na; ++I) {
    if (rhi[I] < 0) throw Error{MHX_ERR_ARG, "AMG: aggregate outside every prolongator row"};
    a.t[kRlo][I] = rlo[I];
    a.t[kRext][I] = rhi[I] - rlo[I] + 1;
    a.t[kSr][I + 1] = a.t[kSr][I] + a.t[kRext][I];
    a.t[kClo][I] = std::max(0, I - rc);
    a.t[kCext][I] = std::min(na - 1, I + rc) - a.t[kClo][I] + 1;
    a.t[kSc][I + 1] = a.t[kSc][I] + a.t[kCext][I];
  }
}

// offset of the first entry of node (x,y,z) in a separable box pattern with
// per-axis extents E and prefix sums S over a lattice of N nodes, nrnc =
// (rows per node) x (columns per box cell)
__device__ __forceinline__ long box_node_base(const int32_t* Sx, const int32_t* Sy, const int32_t* Sz,
                                              const int32_t* Ey, const int32_t* Ez, int Nx, int Ny, int x, int y,
                                              int z, long nrnc) {
  const long Tx = Sx[Nx], Ty = Sy[Ny];
  return nrnc * ((long)Sz[z] * Ty * Tx + (long)Ez[z] * ((long)Sy[y] * Tx + (long)Ey[y] * Sx[x]));
}

// row pointers of a box pattern: lattice N (per-axis extents table e, prefix
// table s), nf rows per node, nf columns per box cell
__global__ void k_box_rowptr(Geo G, int coarse, int e, int s, long nrows, int32_t* __restrict__ rp) {
  const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (row > nrows) return;
  const AxisDev* A = G.ax;
  const int N0 = coarse ? A[0].na : A[0].n, N1 = coarse ? A[1].na : A[1].n;
  const int nf = G.nf;
  if (row == nrows) {
    const int N2 = coarse ? A[2].na : A[2].n;
    rp[row] = (int32_t)((long)nf * nf * A[0].t[s][N0] * (long)A[1].t[s][N1] * (long)A[2].t[s][N2]);
    return;
  }
  const long node = row / nf;
  const int f = (int)(row - node * nf);
  const int x = (int)(node % N0);
  const long yz = node / N0;
  const int y = (int)(yz % N1), z = (int)(yz / N1);
  const long base = box_node_base(A[0].t[s], A[1].t[s], A[2].t[s], A[1].t[e], A[2].t[e], N0, N1, x, y, z, (long)nf * nf);
  rp[row] = (int32_t)(base + (long)f * A[0].t[e][x] * A[1].t[e][y] * A[2].t[e][z] * nf);
}

// lattice coordinates of a node; node numbers fit 32 bits (the CRS indices
// are int32), and 32-bit unsigned division is several times cheaper on the
// GPU than the 64-bit one
__device__ __forceinline__ void node_xyz(long node, int N0, int N1, int& x, int& y, int& z) {
  const unsigned nd = (unsigned)node, n0 = (unsigned)N0, n1 = (unsigned)N1;
  const unsigned yz = nd / n0;
  x = (int)(nd - yz * n0);
  const unsigned zz = yz / n1;
  y = (int)(yz - zz * n1);
  z = (int)zz;
}

// ---------------------------------------------------------------------------
// set-up kernels (FP64)
// ---------------------------------------------------------------------------
__global__ void k_amg_diag(long n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                           const double* __restrict__ A, double* __restrict__ dinv) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = rp[i], hi = rp[i + 1];
  const int end = hi;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ci[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  const double d = (lo < end && ci[lo] == i) ? A[lo] : 0.0;
  dinv[i] = d != 0.0 ? 1.0 / d : 0.0;
}

__global__ void k_amg_ns0(long n, const uint8_t* __restrict__ fixed, double* __restrict__ ns) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ns[i] = fixed && fixed[i] ? 0.0 : 1.0;
}

// power-iteration start vector 0.5 + frac(i * 0.618...), rounded op by op
__global__ void k_amg_start(long n, double* __restrict__ v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = __dmul_rn((double)i, 0.6180339887498949);
  v[i] = __dadd_rn(0.5, __dsub_rn(x, floor(x)));
}


// nrm[(I, f)] = || ns restricted to aggregate I, field f ||
__global__ void k_amg_nrm(Geo G, const double* __restrict__ ns, double* __restrict__ nrm, long ncrows) {
  const long cr = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cr >= ncrows) return;
  const int nf = G.nf;
  const long I = cr / nf;
  const int f = (int)(cr - I * nf);
  int Ix, Iy, Iz;
  node_xyz(I, G.ax[0].na, G.ax[1].na, Ix, Iy, Iz);
  const int Ic[3] = {Ix, Iy, Iz};
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = Ic[d] * G.ax[d].b;
    hi[d] = Ic[d] == G.ax[d].na - 1 ? G.ax[d].n - 1 : lo[d] + G.ax[d].b - 1;
  }
  double s = 0.0;
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const double v = ns[((long)x + (long)G.ax[0].n * ((long)y + (long)G.ax[1].n * z)) * nf + f];
        s = fma(v, v, s);
      }
  nrm[cr] = sqrt(s);
}

// tentative prolongator entry of fine DOF j (its own aggregate and field)
__device__ __forceinline__ double t_entry(const Geo& G, const double* ns, const double* nrm, long j, int* agg3,
                                          int& g) {
  const int nf = G.nf;
  const long q = (long)((unsigned)j / (unsigned)nf);
  g = (int)(j - q * nf);
  int x, y, z;
  node_xyz(q, G.ax[0].n, G.ax[1].n, x, y, z);
  agg3[0] = G.ax[0].t[kAgg][x];
  agg3[1] = G.ax[1].t[kAgg][y];
  agg3[2] = G.ax[2].t[kAgg][z];
  const double v = ns[j];
  if (v == 0.0) return 0.0;
  const long I = agg3[0] + (long)G.ax[0].na * (agg3[1] + (long)G.ax[1].na * agg3[2]);
  const double nv = nrm[I * nf + g];
  return nv > 0.0 ? v / nv : 0.0;
}

// P = (I - omega D^-1 A) T, one thread per fine row; the row's box of at most
// 2x2x2 aggregates x nf fields accumulates in LDS
__global__ __launch_bounds__(kAmgBlk) void k_amg_P(Geo G, long n, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ci, const double* __restrict__ A,
                                                   const double* __restrict__ dinv, const double* __restrict__ ns,
                                                   const double* __restrict__ nrm, const double* __restrict__ lam,
                                                   const int32_t* __restrict__ prp, int32_t* __restrict__ pci,
                                                   double* __restrict__ P, int* __restrict__ err) {
  constexpr int kSlots = 8 * kAmgMaxFields + 1;
  __shared__ double acc_s[kAmgBlk][kSlots];
  const long i = (long)blockIdx.x * kAmgBlk + threadIdx.x;
  if (i >= n) return;
  double* acc = acc_s[threadIdx.x];
  const int nf = G.nf;
  const long p = i / nf;
  int px, py, pz;
  node_xyz(p, G.ax[0].n, G.ax[1].n, px, py, pz);
  const int lo[3] = {G.ax[0].t[kPlo][px], G.ax[1].t[kPlo][py], G.ax[2].t[kPlo][pz]};
  const int ex = G.ax[0].t[kPext][px], ey = G.ax[1].t[kPext][py], ez = G.ax[2].t[kPext][pz];
  const int cnt = ex * ey * ez * nf;
  for (int s = 0; s < cnt; ++s) acc[s] = 0.0;
  auto slot = [&](const int* a3, int g) -> int {
    const int dx = a3[0] - lo[0], dy = a3[1] - lo[1], dz = a3[2] - lo[2];
    if (dx < 0 || dx >= ex || dy < 0 || dy >= ey || dz < 0 || dz >= ez) return -1;
    return ((dz * ey + dy) * ex + dx) * nf + g;
  };
  int a3[3], g;
  const double ti = t_entry(G, ns, nrm, i, a3, g);
  if (ti != 0.0) {
    const int s = slot(a3, g);
    if (s < 0) atomicOr(err, 1);
    else acc[s] += ti;
  }
  const double omega = kAmgDamping / *lam;
  const double sc = -omega * dinv[i];
  for (int k = rp[i]; k < rp[i + 1]; ++k) {
    const double tj = t_entry(G, ns, nrm, ci[k], a3, g);
    if (tj == 0.0) continue;
    const int s = slot(a3, g);
    if (s < 0) {
      atomicOr(err, 1);
      continue;
    }
    acc[s] = fma(sc * A[k], tj, acc[s]);
  }
  const long base = prp[i];
  const int nax = G.ax[0].na, nay = G.ax[1].na;
  for (int s = 0; s < cnt; ++s) {
    const int gg = s % nf, cell = s / nf;
    const int dx = cell % ex, dy = (cell / ex) % ey, dz = cell / (ex * ey);
    const long J = (lo[0] + dx) + (long)nax * ((lo[1] + dy) + (long)nay * (lo[2] + dz));
    pci[base + s] = (int32_t)(J * nf + gg);
    P[base + s] = acc[s];
  }
}

// R = P^T: one workgroup per coarse row (I, f); its columns are the fine DOFs
// of nodes rlo(I) .. rlo(I)+rext(I)-1 per axis, every field
__global__ __launch_bounds__(kAmgBlk) void k_amg_R(Geo G, const int32_t* __restrict__ prp,
                                                   const double* __restrict__ P, const int32_t* __restrict__ rrp,
                                                   int32_t* __restrict__ rci, double* __restrict__ R) {
  const long cr = blockIdx.x;
  const int nf = G.nf;
  const long I = cr / nf;
  const int f = (int)(cr - I * nf);
  int Ix, Iy, Iz;
  node_xyz(I, G.ax[0].na, G.ax[1].na, Ix, Iy, Iz);
  const int rlx = G.ax[0].t[kRlo][Ix], rly = G.ax[1].t[kRlo][Iy], rlz = G.ax[2].t[kRlo][Iz];
  const int rex = G.ax[0].t[kRext][Ix], rey = G.ax[1].t[kRext][Iy];
  const long base = rrp[cr];
  const int cnt = rrp[cr + 1] - rrp[cr];
  for (int e = threadIdx.x; e < cnt; e += kAmgBlk) {
    const int g = e % nf, cell = e / nf;
    const int x = rlx + cell % rex, y = rly + (cell / rex) % rey, z = rlz + cell / (rex * rey);
    const long node = x + (long)G.ax[0].n * (y + (long)G.ax[1].n * z);
    const long i = node * nf + g;
    // position of (I, f) in P row i
    const int ox = Ix - G.ax[0].t[kPlo][x], oy = Iy - G.ax[1].t[kPlo][y], oz = Iz - G.ax[2].t[kPlo][z];
    const int pex = G.ax[0].t[kPext][x], pey = G.ax[1].t[kPext][y];
    rci[base + e] = (int32_t)i;
    R[base + e] = P[prp[i] + ((oz * pey + oy) * pex + ox) * nf + f];
  }
}

// AP(i, (J, g)) in a dense 3x3x3 aggregate box per (fine row i, coarse field
// g).  A group of kApLanes lanes owns one (i, g): lane l takes the row's
// entries l, l + kApLanes, ... (coalesced reads of the CRS row), accumulates
// its own 27-slot box in LDS, and the group then sums the boxes slot by slot
// in lane order, so every set-up gives the same bits.
constexpr int kApLanes = 8;
__global__ __launch_bounds__(kAmgBlk) void k_amg_AP(Geo G, long n, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci, const double* __restrict__ A,
                                                    const int32_t* __restrict__ prp, const double* __restrict__ P,
                                                    double* __restrict__ AP, int* __restrict__ err) {
  __shared__ double acc_s[kAmgBlk][28];
  const int nf = G.nf;
  const int l = threadIdx.x % kApLanes;
  const long t = ((long)blockIdx.x * kAmgBlk + threadIdx.x) / kApLanes;
  const bool active = t < n * nf;
  double* acc = acc_s[threadIdx.x];
  for (int s = 0; s < 27; ++s) acc[s] = 0.0;
  const long i = active ? t / nf : 0;
  const int g = active ? (int)(t - i * nf) : 0;
  if (active) {
    int px, py, pz;
    node_xyz((long)((unsigned)i / (unsigned)nf), G.ax[0].n, G.ax[1].n, px, py, pz);
    const int alo[3] = {G.ax[0].t[kAplo][px], G.ax[1].t[kAplo][py], G.ax[2].t[kAplo][pz]};
    const int aex[3] = {G.ax[0].t[kApext][px], G.ax[1].t[kApext][py], G.ax[2].t[kApext][pz]};
    for (int k = rp[i] + l; k < rp[i + 1]; k += kApLanes) {
      const long j = ci[k];
      const double a = A[k];
      int qx, qy, qz;
      node_xyz((long)((unsigned)j / (unsigned)nf), G.ax[0].n, G.ax[1].n, qx, qy, qz);
      const int plx = G.ax[0].t[kPlo][qx], ply = G.ax[1].t[kPlo][qy], plz = G.ax[2].t[kPlo][qz];
      const int pex = G.ax[0].t[kPext][qx], pey = G.ax[1].t[kPext][qy], pez = G.ax[2].t[kPext][qz];
      const long pb = prp[j];
      for (int dz = 0; dz < pez; ++dz)
        for (int dy = 0; dy < pey; ++dy)
          for (int dx = 0; dx < pex; ++dx) {
            const int sx = plx + dx - alo[0], sy = ply + dy - alo[1], sz = plz + dz - alo[2];
            if (sx < 0 || sx >= aex[0] || sy < 0 || sy >= aex[1] || sz < 0 || sz >= aex[2]) {
              atomicOr(err, 2);
              continue;
            }
            const double pv = P[pb + ((dz * pey + dy) * pex + dx) * nf + g];
            double& r = acc[(sz * 3 + sy) * 3 + sx];
            r = fma(a, pv, r);
          }
    }
  }
  __syncthreads();
  if (!active) return;
  const int g0 = threadIdx.x - l;  // first lane of the group
  for (int s = l; s < 27; s += kApLanes) {
    double v = 0.0;
    for (int q = 0; q < kApLanes; ++q) v += acc_s[g0 + q][s];
    AP[t * 27 + s] = v;
  }
}

// A_c = R (A P): one workgroup per coarse row (I, f), each thread owning up to
// 4 entries (J, g) of the row; the R row streams through LDS in chunks and
// every entry sums in R's order
__global__ __launch_bounds__(kAmgBlk) void k_amg_Ac(Geo G, const int32_t* __restrict__ rrp,
                                                    const int32_t* __restrict__ rci, const double* __restrict__ R,
                                                    const double* __restrict__ AP, const double* __restrict__ nrm,
                                                    const int32_t* __restrict__ arp, int32_t* __restrict__ aci,
                                                    double* __restrict__ Ac) {
  constexpr int kPer = 4;
  __shared__ double s_val[kAmgBlk];
  __shared__ long s_row[kAmgBlk];
  __shared__ int s_lo[kAmgBlk][3], s_ex[kAmgBlk][3];
  const long cr = blockIdx.x;
  const int nf = G.nf;
  const long I = cr / nf;
  int Ix, Iy, Iz;
  node_xyz(I, G.ax[0].na, G.ax[1].na, Ix, Iy, Iz);
  const int clx = G.ax[0].t[kClo][Ix], cly = G.ax[1].t[kClo][Iy], clz = G.ax[2].t[kClo][Iz];
  const int cex = G.ax[0].t[kCext][Ix], cey = G.ax[1].t[kCext][Iy];
  const long base = arp[cr];
  const int cnt = arp[cr + 1] - arp[cr];
  int J3[kPer][3], gg[kPer];
  double sum[kPer];
#pragma unroll
  for (int m = 0; m < kPer; ++m) {
    sum[m] = 0.0;
    const int e = threadIdx.x + m * kAmgBlk;
    const int g = e % nf, cell = e / nf;
    gg[m] = g;
    J3[m][0] = clx + cell % cex;
    J3[m][1] = cly + (cell / cex) % cey;
    J3[m][2] = clz + cell / (cex * cey);
  }
  const int rb = rrp[cr], re = rrp[cr + 1];
  for (int c0 = rb; c0 < re; c0 += kAmgBlk) {
    const int k = c0 + threadIdx.x;
    if (k < re) {
      const long i = rci[k];
      int px, py, pz;
      node_xyz((long)((unsigned)i / (unsigned)nf), G.ax[0].n, G.ax[1].n, px, py, pz);
      s_val[threadIdx.x] = R[k];
      s_row[threadIdx.x] = i;
      s_lo[threadIdx.x][0] = G.ax[0].t[kAplo][px];
      s_lo[threadIdx.x][1] = G.ax[1].t[kAplo][py];
      s_lo[threadIdx.x][2] = G.ax[2].t[kAplo][pz];
      s_ex[threadIdx.x][0] = G.ax[0].t[kApext][px];
      s_ex[threadIdx.x][1] = G.ax[1].t[kApext][py];
      s_ex[threadIdx.x][2] = G.ax[2].t[kApext][pz];
    }
    __syncthreads();
    const int nk = min(kAmgBlk, re - c0);
    for (int c = 0; c < nk; ++c) {
      const double rv = s_val[c];
      const long i = s_row[c];
#pragma unroll
      for (int m = 0; m < kPer; ++m) {
        if (threadIdx.x + m * kAmgBlk >= cnt) break;
        const int sx = J3[m][0] - s_lo[c][0], sy = J3[m][1] - s_lo[c][1], sz = J3[m][2] - s_lo[c][2];
        if (sx < 0 || sx >= s_ex[c][0] || sy < 0 || sy >= s_ex[c][1] || sz < 0 || sz >= s_ex[c][2]) continue;
        sum[m] = fma(rv, AP[(i * nf + gg[m]) * 27 + (sz * 3 + sy) * 3 + sx], sum[m]);
      }
    }
    __syncthreads();
  }
  const int nax = G.ax[0].na, nay = G.ax[1].na;
#pragma unroll
  for (int m = 0; m < kPer; ++m) {
    const int e = threadIdx.x + m * kAmgBlk;
    if (e >= cnt) break;
    const long col = (J3[m][0] + (long)nax * (J3[m][1] + (long)nay * J3[m][2])) * nf + gg[m];
    aci[base + e] = (int32_t)col;
    // an aggregate whose near-null space vanished (all its nodes strong
    // Dirichlet) has an empty prolongator column: an identity row keeps A_c
    // invertible without coupling it to anything
    Ac[base + e] = (col == cr && nrm[cr] == 0.0) ? 1.0 : sum[m];
  }
}

// dense [A | I] of the coarsest level from its CRS
__global__ void k_amg_dense(long n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                            const double* __restrict__ A, double* __restrict__ M) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* row = M + i * 2 * n;
  for (long j = 0; j < 2 * n; ++j) row[j] = (j - n == i) ? 1.0 : 0.0;
  for (int k = rp[i]; k < rp[i + 1]; ++k) row[ci[k]] += A[k];
}

// Gauss-Jordan with partial pivoting on [A | I] (n x 2n, row-major) in one
// workgroup; afterwards the right half holds A^-1.  F: n factors of scratch.
__global__ __launch_bounds__(1024) void k_amg_gauss_jordan(int n, double* __restrict__ M, double* __restrict__ F,
                                                           int* __restrict__ err) {
  __shared__ double sv[1024];
  __shared__ int si[1024];
  const int t = threadIdx.x, w = 2 * n;
  for (int col = 0; col < n; ++col) {
    double best = -1.0;
    int bi = col;
    for (int r = col + t; r < n; r += 1024) {
      const double v = fabs(M[(long)r * w + col]);
      if (v > best) {
        best = v;
        bi = r;
      }
    }
    sv[t] = best;
    si[t] = bi;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
      if (t < s && (sv[t + s] > sv[t] || (sv[t + s] == sv[t] && si[t + s] < si[t]))) {
        sv[t] = sv[t + s];
        si[t] = si[t + s];
      }
      __syncthreads();
    }
    const int pr = si[0];
    const double pmax = sv[0];
    __syncthreads();
    if (!(pmax > 0.0)) {
      if (t == 0) atomicOr(err, 4);
      return;  // uniform across the workgroup
    }
    if (pr != col)
      for (int j = t; j < w; j += 1024) {
        const double a = M[(long)pr * w + j];
        M[(long)pr * w + j] = M[(long)col * w + j];
        M[(long)col * w + j] = a;
      }
    __syncthreads();
    const double inv = 1.0 / M[(long)col * w + col];
    __syncthreads();
    for (int j = t; j < w; j += 1024) M[(long)col * w + j] *= inv;
    for (int r = t; r < n; r += 1024) F[r] = r == col ? 0.0 : M[(long)r * w + col];
    __syncthreads();
    // thread t owns column t (w <= 1024); left-half columns before the pivot
    // are already unit vectors with a zero in the pivot row, so they stay
    if (t < w && (t >= col)) {
      const double pc = M[(long)col * w + t];
      if (pc != 0.0)
        for (int r = 0; r < n; ++r)
          if (r != col) M[(long)r * w + t] = fma(-F[r], pc, M[(long)r * w + t]);
    }
    __syncthreads();
  }
}

template <class T>
__global__ void k_amg_extract_inv(long n, const double* __restrict__ M, T* __restrict__ out) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const long r = idx / n, c = idx - r * n;
  out[idx] = (T)M[r * 2 * n + n + c];
}

template <class TI, class TO>
__global__ void k_amg_cast(long n, const TI* __restrict__ a, TO* __restrict__ b) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (TO)a[i];
}

// nodal block inverses: the BS x BS diagonal block of node k (DOFs k*BS ..
// k*BS+BS-1, consecutive columns in every sorted row), Gauss-Jordan with
// partial pivoting in FP64; a zero pivot leaves a zero block (no smoothing of
// that node) instead of Inf
template <int BS>
__global__ void k_amg_block_inv(long nblk, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                const double* __restrict__ A, double* __restrict__ out) {
  const long blk = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (blk >= nblk) return;
  double M[BS][BS], I[BS][BS];
  const int c0 = (int)(blk * BS);
  for (int i = 0; i < BS; ++i) {
    int lo = rp[c0 + i], hi = rp[c0 + i + 1];
    const int end = hi;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (ci[mid] < c0) lo = mid + 1;
      else hi = mid;
    }
    int k = lo;
    for (int j = 0; j < BS; ++j) {
      const bool hit = k < end && ci[k] == c0 + j;
      M[i][j] = hit ? A[k] : 0.0;
      k += hit;
      I[i][j] = i == j ? 1.0 : 0.0;
    }
  }
  bool ok = true;
  for (int col = 0; col < BS; ++col) {
    int piv = col;
    double best = fabs(M[col][col]);
    for (int r = col + 1; r < BS; ++r)
      if (fabs(M[r][col]) > best) {
        best = fabs(M[r][col]);
        piv = r;
      }
    if (!(best > 0.0)) {
      ok = false;
      break;
    }
    if (piv != col)
      for (int j = 0; j < BS; ++j) {
        double t = M[col][j]; M[col][j] = M[piv][j]; M[piv][j] = t;
        t = I[col][j]; I[col][j] = I[piv][j]; I[piv][j] = t;
      }
    const double d = 1.0 / M[col][col];
    for (int j = 0; j < BS; ++j) {
      M[col][j] *= d;
      I[col][j] *= d;
    }
    for (int r = 0; r < BS; ++r) {
      if (r == col) continue;
      const double f = M[r][col];
      for (int j = 0; j < BS; ++j) {
        M[r][j] = fma(-f, M[col][j], M[r][j]);
        I[r][j] = fma(-f, I[col][j], I[r][j]);
      }
    }
  }
  for (int i = 0; i < BS; ++i)
    for (int j = 0; j < BS; ++j) out[blk * BS * BS + i * BS + j] = ok ? I[i][j] : 0.0;
}

// xout = (zero_x ? 0 : xin) + omega Binv r, one DOF per thread (in place allowed)
template <class T>
__global__ void k_amg_bj(long n, int nf, const T* __restrict__ binv, const T* __restrict__ r, const T* xin, T* xout,
                         double omega, int zero_x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long node = (long)((unsigned)i / (unsigned)nf);  // rows fit 32 bits
  const int f = (int)(i - node * nf);
  const T* row = binv + node * nf * nf + f * nf;
  const T* rb = r + node * nf;
  double s = 0.0;
  for (int g = 0; g < nf; ++g) s = fma((double)row[g], (double)rb[g], s);
  xout[i] = (T)((zero_x ? 0.0 : (double)xin[i]) + omega * s);
}

// ---------------------------------------------------------------------------
// V-cycle kernels (T = double or float; sums in FP64)
// ---------------------------------------------------------------------------
template <int LANES, class T, class CI>
__device__ __forceinline__ double row_dot(const int32_t* __restrict__ rp, CI ci, const T* __restrict__ A,
                                          const T* __restrict__ x, long row, int l) {
  const int e = rp[row + 1];
  double acc = 0.0;
  int k = rp[row] + l;
  for (; k + 3 * LANES < e; k += 4 * LANES) {
    T av[4];
    int cv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      av[u] = A[k + u * LANES];
      cv[u] = ci(k + u * LANES);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) acc = fma((double)av[u], (double)x[cv[u]], acc);
  }
  for (; k < e; k += LANES) acc = fma((double)A[k], (double)x[ci(k)], acc);
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
  return acc;
}

// one Chebyshev step: r = b - A xin (A xin skipped when zero_x),
// d = alpha d + beta D^-1 r, xout = xin + d (xin = 0 when zero_x)
// LAT: the level-0 operator in the context's lattice pattern, columns from its
// closed form instead of ci (same entries, same order, same sums)
template <bool LAT>
__device__ __forceinline__ auto level_cols(const BoxMesh& m, const int32_t* __restrict__ rp,
                                           const int32_t* __restrict__ ci, long row) {
  if constexpr (LAT) return LatCols{lat_row(m, row), rp[row]};
  else return CsrCols{ci};
}

// w = s * D^-1 A v  (s = *sp, on the device: the scale applied to v, so v
// need not be normalised in place)
template <int LANES, bool LAT>
__global__ __launch_bounds__(256) void k_amg_dinv_spmv(BoxMesh m, long n, const int32_t* __restrict__ rp,
                                                       const int32_t* __restrict__ ci, const double* __restrict__ A,
                                                       const double* __restrict__ dinv, const double* __restrict__ sp,
                                                       const double* __restrict__ v, double* __restrict__ w) {
  const long row = ((long)blockIdx.x * blockDim.x + threadIdx.x) / LANES;
  const int l = threadIdx.x % LANES;
  if (row >= n) return;
  const double acc = row_dot<LANES, double>(rp, level_cols<LAT>(m, rp, ci, row), A, v, row, l);
  if (l == 0) w[row] = *sp * dinv[row] * acc;
}

// scalar steps of the power iteration, one thread; st = (s, lambda, stopped).
// phase 0: s = 1 / sqrt(v.v); 1: lambda = s (v.w), then s = 1 / |w| unless
// |w| is zero or not finite, which stops the iteration (later steps leave
// lambda alone); 2: lambda = 1 where it is not a finite positive number.
__global__ void k_amg_power(int phase, const double* __restrict__ d0, const double* __restrict__ d1,
                            double* __restrict__ st) {
  if (phase == 0) {
    st[0] = 1.0 / sqrt(*d0);
    st[1] = 1.0;
    st[2] = 0.0;
  } else if (phase == 1) {
    if (st[2] != 0.0) return;
    st[1] = st[0] * *d0;
    const double wn = sqrt(*d1);
    if (!(wn > 0.0) || !isfinite(wn)) st[2] = 1.0;
    else st[0] = 1.0 / wn;
  } else {
    const double lam = st[1];
    st[1] = (isfinite(lam) && lam > 0.0) ? lam : 1.0;
  }
}

template <int LANES, bool LAT, class T>
__global__ __launch_bounds__(256) void k_amg_cheb(BoxMesh m, long n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ci, const T* __restrict__ A,
                                                  const T* __restrict__ b, const T* __restrict__ xin,
                                                  T* __restrict__ xout, T* __restrict__ d,
                                                  const T* __restrict__ dinv, double alpha, double beta, int zero_x) {
  const long row = ((long)blockIdx.x * blockDim.x + threadIdx.x) / LANES;
  const int l = threadIdx.x % LANES;
  if (row >= n) return;
  const double ax = zero_x ? 0.0 : row_dot<LANES, T>(rp, level_cols<LAT>(m, rp, ci, row), A, xin, row, l);
  if (l == 0) {
    const double r = (double)b[row] - ax;
    const double dn = (alpha != 0.0 ? alpha * (double)d[row] : 0.0) + beta * (double)dinv[row] * r;
    d[row] = (T)dn;
    xout[row] = (T)((zero_x ? 0.0 : (double)xin[row]) + dn);
  }
}

// r = b - A x
template <int LANES, bool LAT, class T>
__global__ __launch_bounds__(256) void k_amg_resid(BoxMesh m, long n, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ci, const T* __restrict__ A,
                                                   const T* __restrict__ b, const T* __restrict__ x,
                                                   T* __restrict__ r) {
  const long row = ((long)blockIdx.x * blockDim.x + threadIdx.x) / LANES;
  const int l = threadIdx.x % LANES;
  if (row >= n) return;
  const double ax = row_dot<LANES, T>(rp, level_cols<LAT>(m, rp, ci, row), A, x, row, l);
  if (l == 0) r[row] = (T)((double)b[row] - ax);
}

// One colour of a multicolour nodal block Gauss-Seidel sweep: for every node
// of the colour, x_node += Binv_node (b - A x)_node.  Colour (cx, cy, cz) of
// stride s holds the nodes whose lattice coordinates are congruent to it
// modulo s; with s = reach + 1 no two nodes of a colour share a matrix entry,
// so the in-place update is race free.  A group of NFP * LANES lanes (NFP = nf
// rounded up to a power of two, groups never straddle a wavefront) owns a
// node, LANES lanes per row; the rows' residuals meet through wavefront
// shuffles, which every lane executes.
template <int LANES, int NFP, class T>
__global__ __launch_bounds__(256) void k_amg_gs(long ncol, int nf, int s, int cx, int cy, int cz, long cntx, long cnty,
                                                long Nx, long Ny, const int32_t* __restrict__ rp,
                                                const int32_t* __restrict__ ci, const T* __restrict__ A,
                                                const T* __restrict__ binv, const T* __restrict__ b, T* x) {
  static_assert(NFP * LANES <= 64 && 64 % (NFP * LANES) == 0, "node group must tile a wavefront");
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long t = gid / (NFP * LANES);
  const int within = (int)(gid % (NFP * LANES));
  const int f = within / LANES, l = within % LANES;
  const bool active = t < ncol && f < nf;
  long node = 0, row = 0;
  double res = 0.0;
  if (active) {
    // node numbers fit 32 bits: 32-bit unsigned divisions
    const unsigned tu = (unsigned)t, q = tu / (unsigned)cntx, qz = q / (unsigned)cnty;
    const long a = (long)(tu - q * (unsigned)cntx), qy = (long)(q - qz * (unsigned)cnty);
    node = (cx + s * a) + Nx * ((cy + s * qy) + Ny * (cz + s * (long)qz));
    row = node * nf + f;
    res = (double)b[row] - row_dot<LANES, T>(rp, CsrCols{ci}, A, x, row, l);
  }
  const int base = (int)(threadIdx.x & 63) - within;
  double upd = 0.0;
#pragma unroll
  for (int g = 0; g < NFP; ++g) {
    const double rg = __shfl(res, base + g * LANES, 64);
    if (active && g < nf) upd = fma((double)binv[(node * nf + f) * nf + g], rg, upd);
  }
  if (active && l == 0) x[row] = (T)((double)x[row] + upd);
}

template <class T>
__global__ void k_amg_fill0(long n, T* x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (T)0;
}

// y = M x (add = 0) or y += M x (add = 1)
template <int LANES, class T>
__global__ __launch_bounds__(256) void k_amg_spmv(long n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ci, const T* __restrict__ M,
                                                  const T* __restrict__ x, T* __restrict__ y, int add) {
  const long row = ((long)blockIdx.x * blockDim.x + threadIdx.x) / LANES;
  const int l = threadIdx.x % LANES;
  if (row >= n) return;
  const double v = row_dot<LANES, T>(rp, CsrCols{ci}, M, x, row, l);
  if (l == 0) y[row] = (T)(add ? (double)y[row] + v : v);
}

// x = Cinv b, one wavefront per row of the dense coarsest inverse.  The
// inverse stays FP64 in both precisions: in FP32 its rounding error would be
// amplified by the condition number of the coarsest operator.
template <class T>
__global__ __launch_bounds__(256) void k_amg_gemv(long n, const double* __restrict__ Cinv, const T* __restrict__ b,
                                                  T* __restrict__ x) {
  const long row = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int l = threadIdx.x % 64;
  if (row >= n) return;
  double acc = 0.0;
  for (long j = l; j < n; j += 64) acc = fma((double)Cinv[row * n + j], (double)b[j], acc);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (l == 0) x[row] = (T)acc;
}

// ---------------------------------------------------------------------------
// hierarchy
// ---------------------------------------------------------------------------
template <class T>
struct AmgVals {
  DevBuf<T> A, dinv, P, R, cinv, binv;
  DevBuf<T> x, t, b, r, d;  // level work vectors
};

struct AmgLevel {
  long n = 0;            // rows (nodes x fields)
  int N[3] = {1, 1, 1};  // node lattice
  int reach = 1;
  int lanes = 4, plane = 4, rlane = 4;  // lanes per row: A, P, R
  const int32_t* rp = nullptr;  // level 0: the context's pattern
  const int32_t* ci = nullptr;
  DevBuf<int32_t> rp_own, ci_own;
  DevBuf<double> A, dinv, ns;
  DevBuf<int32_t> prp, pci, rrp, rci, tab;
  DevBuf<double> P, R;
  DevBuf<double> binv;  // nodal nf x nf diagonal-block inverses (block-Jacobi smoother)
  bool lat = false;     // level 0 in the context's lattice pattern: closed-form columns
  double lambda = 1.0;
  AmgVals<float> f32;
  AmgVals<double> f64;  // FP64 work vectors (values live in A/P/R above)
};

struct Amg {
  std::vector<std::unique_ptr<AmgLevel>> lv;
  DevBuf<double> cinv;    // dense inverse of the coarsest operator (FP64 in both precisions)
  DevBuf<double> scratch;  // set-up temporaries
  DevBuf<int> err;
  // power iterations on the device: two inner-product slots, then per level
  // (s, lambda, stopped, -)
  DevBuf<double> pw;
  const double* A0 = nullptr;  // level-0 operator of the last set-up
  bool fp32 = false;
  int smoother = kSmootherChebyshev;
  int sweeps = kBjSweeps;   // block-Jacobi sweeps before and after the coarse correction
  double omega = kBjOmega;  // block-Jacobi damping
  int nf = 1;
  int64_t setup_count = 0;
  int64_t reuse_count = 0;  // set-ups answered by refreshing the fine level only
};

static int lanes_for(double avg) {
  int l = 4;
  while (l < 64 && l * 14 < avg) l *= 2;
  return l;
}

static dim3 g1(long n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

template <class K>
static void by_lanes(int lanes, K&& k) {
  switch (lanes) {
    case 4: k(std::integral_constant<int, 4>{}); break;
    case 8: k(std::integral_constant<int, 8>{}); break;
    case 16: k(std::integral_constant<int, 16>{}); break;
    case 32: k(std::integral_constant<int, 32>{}); break;
    default: k(std::integral_constant<int, 64>{}); break;
  }
}

template <class F>
static void by_lat(bool lat, F&& f) {
  if (lat) f(std::true_type{});
  else f(std::false_type{});
}

static Geo upload_geo(Ctx& c, AmgLevel& L, const AxisHost* ax, int nf) {
  size_t total = 0;
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < kNumTables; ++k) total += ax[d].t[k].size();
  L.tab.alloc(total);
  std::vector<int32_t> h(total);
  Geo G{};
  size_t off = 0;
  for (int d = 0; d < 3; ++d) {
    G.ax[d].n = ax[d].n;
    G.ax[d].na = ax[d].na;
    G.ax[d].b = ax[d].b;
    for (int k = 0; k < kNumTables; ++k) {
      std::copy(ax[d].t[k].begin(), ax[d].t[k].end(), h.begin() + off);
      G.ax[d].t[k] = L.tab.p + off;
      off += ax[d].t[k].size();
    }
  }
  MHX_HIP(hipMemcpyAsync(L.tab.p, h.data(), total * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
  G.nf = nf;
  return G;
}

// the set-up's one host synchronisation: its error bits (1, 2: an entry of
// the operator outside the lattice stencil while forming P or R; 4: singular
// coarsest operator) and the lambda of each of the first n_smoothed levels
static void finish_setup(Ctx& c, Amg& H, int n_smoothed) {
  int e = 0;
  std::vector<double> st(2 + 4 * (size_t)n_smoothed);
  MHX_HIP(hipMemcpyAsync(&e, H.err.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  if (n_smoothed > 0)
    MHX_HIP(hipMemcpyAsync(st.data(), H.pw.p, st.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MHX_HIP(hipStreamSynchronize(c.stream));
  for (int l = 0; l < n_smoothed; ++l) H.lv[l]->lambda = st[2 + 4 * l + 1];
  if (e & 4) throw Error{MHX_ERR_NONFINITE, "AMG: singular coarsest operator"};
  if (e) throw Error{MHX_ERR_ARG, "AMG: matrix entries outside the lattice stencil (prolongator)"};
}

// lambda ~ rho(D^-1 A) by power iteration from a fixed start vector,
// entirely on the stream: the scale s and the estimate stay on the device
// (st = (s, lambda, stopped)), and the host reads every level's lambda once,
// at the end of the set-up.  The steps are the host form's, in the same
// floating-point operations: lambda = s (v . D^-1 A s v), s = 1 / |w|.
static void power_estimate(Ctx& c, AmgLevel& L, const double* A, double* v, double* w, double* slots,
                           double* st) {
  const long n = L.n;
  hipLaunchKernelGGL(k_amg_start, g1(n), dim3(256), 0, c.stream, n, v);
  MHX_HIP(hipGetLastError());
  dot_dev(c, v, v, n, slots);
  hipLaunchKernelGGL(k_amg_power, dim3(1), dim3(1), 0, c.stream, 0, slots, slots, st);
  for (int it = 0; it < kAmgPowerIters; ++it) {
    by_lanes(L.lanes, [&](auto LN) {
      by_lat(L.lat, [&](auto LT) {
        hipLaunchKernelGGL((k_amg_dinv_spmv<decltype(LN)::value, decltype(LT)::value>), g1(n * LN), dim3(256), 0,
                           c.stream, c.mesh, n, L.rp, L.ci, A, L.dinv.p, st, v, w);
      });
    });
    MHX_HIP(hipGetLastError());
    dot_dev(c, v, w, n, slots);
    dot_dev(c, w, w, n, slots + 1);
    hipLaunchKernelGGL(k_amg_power, dim3(1), dim3(1), 0, c.stream, 1, slots, slots + 1, st);
    std::swap(v, w);
  }
  hipLaunchKernelGGL(k_amg_power, dim3(1), dim3(1), 0, c.stream, 2, slots, slots, st);
  MHX_HIP(hipGetLastError());
}

void amg_setup(Ctx& c, const double* A, bool fp32) {
  ScopedPhase ph(c.prof, "LinearAlgebraInterface::buildAMGPreconditioner", c.stream);
  if (!c.amg) c.amg = std::make_shared<Amg>();
  Amg& H = *c.amg;
  const int nf = c.nfield;
  MHX_REQUIRE(nf >= 1 && nf <= kAmgMaxFields, MHX_ERR_UNSUPPORTED, "AMG: at most 4 unknowns per node");
  // automatic choice: Chebyshev on the point diagonal for one unknown per
  // node, nodal block Jacobi for coupled fields
  const int smoother = c.amg_smoother >= 0 ? c.amg_smoother : (nf > 1 ? kSmootherBlockJacobi : kSmootherChebyshev);
  // reuse (MueLu "reuse: full" in kind): keep the coarse levels of the last
  // set-up and refresh only what the fine-level smoother reads from the new
  // operator (its diagonal, or its nodal block inverses); the fine operator
  // itself is the new one, passed to every V-cycle
  if (c.amg_reuse && H.setup_count > 0 && H.fp32 == fp32 && H.nf == nf && H.smoother == smoother &&
      !H.lv.empty() && H.lv[0]->n == c.mesh.ndof) {
    AmgLevel& L0 = *H.lv[0];
    H.A0 = A;
    // cycle parameters are read per V-cycle: take the current options
    H.sweeps = c.amg_sweeps;
    H.omega = c.amg_omega_pct * 0.01;
    hipLaunchKernelGGL(k_amg_diag, g1(L0.n), dim3(256), 0, c.stream, L0.n, L0.rp, L0.ci, A, L0.dinv.p);
    if (H.smoother != kSmootherChebyshev && H.lv.size() > 1) {
      const long nblk = L0.n / nf;
      switch (nf) {
        case 1: hipLaunchKernelGGL(k_amg_block_inv<1>, g1(nblk), dim3(256), 0, c.stream, nblk, L0.rp, L0.ci, A, L0.binv.p); break;
        case 2: hipLaunchKernelGGL(k_amg_block_inv<2>, g1(nblk), dim3(256), 0, c.stream, nblk, L0.rp, L0.ci, A, L0.binv.p); break;
        case 3: hipLaunchKernelGGL(k_amg_block_inv<3>, g1(nblk), dim3(256), 0, c.stream, nblk, L0.rp, L0.ci, A, L0.binv.p); break;
        default: hipLaunchKernelGGL(k_amg_block_inv<4>, g1(nblk), dim3(256), 0, c.stream, nblk, L0.rp, L0.ci, A, L0.binv.p); break;
      }
    }
    if (fp32) {
      hipLaunchKernelGGL((k_amg_cast<double, float>), g1(L0.n), dim3(256), 0, c.stream, L0.n, L0.dinv.p,
                         L0.f32.dinv.p);
      if (H.smoother != kSmootherChebyshev && H.lv.size() > 1)
        hipLaunchKernelGGL((k_amg_cast<double, float>), g1((long)L0.binv.n), dim3(256), 0, c.stream,
                           (long)L0.binv.n, L0.binv.p, L0.f32.binv.p);
    }
    MHX_HIP(hipGetLastError());
    ++H.reuse_count;
    return;
  }
  H.err.alloc(1);
  H.pw.alloc(2 + 4 * kAmgMaxLevels);
  MHX_HIP(hipMemsetAsync(H.err.p, 0, sizeof(int), c.stream));
  H.A0 = A;
  H.fp32 = fp32;
  H.smoother = smoother;
  H.sweeps = c.amg_sweeps;
  H.omega = c.amg_omega_pct * 0.01;
  H.nf = nf;
  ++H.setup_count;

  // level 0: the context's lattice and pattern
  int nlev = 0;
  auto next_level = [&]() -> AmgLevel& {
    if ((int)H.lv.size() <= nlev) H.lv.emplace_back(new AmgLevel());
    return *H.lv[nlev++];
  };
  AmgLevel* L = &next_level();
  L->n = c.mesh.ndof;
  for (int d = 0; d < 3; ++d) L->N[d] = d < c.mesh.dim ? c.mesh.nn[d] : 1;
  L->reach = c.mesh.p;
  L->rp = c.rowptr.p;
  L->ci = c.colidx.p;
  L->lanes = c.spmv_lanes;
  L->lat = c.lattice_cols != 0;
  L->dinv.alloc(L->n);
  L->ns.alloc(L->n);
  hipLaunchKernelGGL(k_amg_diag, g1(L->n), dim3(256), 0, c.stream, L->n, L->rp, L->ci, A, L->dinv.p);
  hipLaunchKernelGGL(k_amg_ns0, g1(L->n), dim3(256), 0, c.stream, L->n, c.fixed.p, L->ns.p);
  MHX_HIP(hipGetLastError());
  const double* Acur = A;

  while (L->n > kAmgCoarseMax && nlev < kAmgMaxLevels) {
    const int r = std::max(1, L->reach);
    const int b = 2 * r + 1;
    const int rc = (3 * r - 1) / b + 1;
    AxisHost ax[3];
    for (int d = 0; d < 3; ++d) make_axis(ax[d], L->N[d], r, rc);
    const long ncn = (long)ax[0].na * ax[1].na * ax[2].na;
    const long nc = ncn * nf;
    if (nc >= L->n) break;  // no coarsening left
    Geo G = upload_geo(c, *L, ax, nf);
    const long n = L->n;
    // lambda (on the device; omega = damping / lambda is formed by k_amg_P)
    H.scratch.alloc(std::max<size_t>(H.scratch.n, (size_t)2 * n));
    double* st = H.pw.p + 2 + 4 * (nlev - 1);
    power_estimate(c, *L, Acur, H.scratch.p, H.scratch.p + n, H.pw.p, st);
    // next level's near-null space = aggregate norms
    AmgLevel& C = next_level();
    L = H.lv[nlev - 2].get();  // the vector may have grown
    C.n = nc;
    for (int d = 0; d < 3; ++d) C.N[d] = ax[d].na;
    C.reach = rc;
    C.ns.alloc(nc);
    hipLaunchKernelGGL(k_amg_nrm, g1(nc), dim3(256), 0, c.stream, G, L->ns.p, C.ns.p, nc);
    // P
    const long p_nnz = (long)nf * nf * ax[0].t[kSp][ax[0].n] * (long)ax[1].t[kSp][ax[1].n] * ax[2].t[kSp][ax[2].n];
    const long r_nnz = (long)nf * nf * ax[0].t[kSr][ax[0].na] * (long)ax[1].t[kSr][ax[1].na] * ax[2].t[kSr][ax[2].na];
    const long a_nnz = (long)nf * nf * ax[0].t[kSc][ax[0].na] * (long)ax[1].t[kSc][ax[1].na] * ax[2].t[kSc][ax[2].na];
    MHX_REQUIRE(p_nnz < INT_MAX && r_nnz < INT_MAX && a_nnz < INT_MAX, MHX_ERR_UNSUPPORTED,
                "AMG: transfer operator exceeds 32-bit indexing");
    L->prp.alloc(n + 1);
    L->pci.alloc(p_nnz);
    L->P.alloc(p_nnz);
    hipLaunchKernelGGL(k_box_rowptr, g1(n + 1), dim3(256), 0, c.stream, G, 0, (int)kPext, (int)kSp, n, L->prp.p);
    hipLaunchKernelGGL(k_amg_P, g1(n, kAmgBlk), dim3(kAmgBlk), 0, c.stream, G, n, L->rp, L->ci, Acur, L->dinv.p,
                       L->ns.p, C.ns.p, st + 1, L->prp.p, L->pci.p, L->P.p, H.err.p);
    // R = P^T
    L->rrp.alloc(nc + 1);
    L->rci.alloc(r_nnz);
    L->R.alloc(r_nnz);
    hipLaunchKernelGGL(k_box_rowptr, g1(nc + 1), dim3(256), 0, c.stream, G, 1, (int)kRext, (int)kSr, nc, L->rrp.p);
    hipLaunchKernelGGL(k_amg_R, dim3((unsigned)nc), dim3(kAmgBlk), 0, c.stream, G, L->prp.p, L->P.p, L->rrp.p,
                       L->rci.p, L->R.p);
    // A P (dense 27-slot boxes), then A_c = R (A P)
    H.scratch.alloc(std::max<size_t>(H.scratch.n, (size_t)n * nf * 27));
    hipLaunchKernelGGL(k_amg_AP, g1(n * nf * kApLanes, kAmgBlk), dim3(kAmgBlk), 0, c.stream, G, n, L->rp, L->ci, Acur,
                       L->prp.p, L->P.p, H.scratch.p, H.err.p);
    C.rp_own.alloc(nc + 1);
    C.ci_own.alloc(a_nnz);
    C.A.alloc(a_nnz);
    hipLaunchKernelGGL(k_box_rowptr, g1(nc + 1), dim3(256), 0, c.stream, G, 1, (int)kCext, (int)kSc, nc,
                       C.rp_own.p);
    hipLaunchKernelGGL(k_amg_Ac, dim3((unsigned)nc), dim3(kAmgBlk), 0, c.stream, G, L->rrp.p, L->rci.p, L->R.p,
                       H.scratch.p, C.ns.p, C.rp_own.p, C.ci_own.p, C.A.p);
    MHX_HIP(hipGetLastError());
    L->plane = lanes_for((double)p_nnz / (double)n);
    L->rlane = lanes_for((double)r_nnz / (double)nc);
    C.rp = C.rp_own.p;
    C.ci = C.ci_own.p;
    C.lanes = lanes_for((double)a_nnz / (double)nc);
    C.dinv.alloc(nc);
    hipLaunchKernelGGL(k_amg_diag, g1(nc), dim3(256), 0, c.stream, nc, C.rp, C.ci, C.A.p, C.dinv.p);
    MHX_HIP(hipGetLastError());
    Acur = C.A.p;
    L = &C;
  }
  H.lv.resize(nlev);
  // coarsest: dense inverse
  {
    AmgLevel& Z = *H.lv.back();
    const long n = Z.n;
    MHX_REQUIRE(n <= kAmgDenseMax, MHX_ERR_UNSUPPORTED, "AMG: coarsest level too large for the dense solve");
    H.scratch.alloc(std::max<size_t>(H.scratch.n, (size_t)2 * n * n + n));
    hipLaunchKernelGGL(k_amg_dense, g1(n), dim3(256), 0, c.stream, n, Z.rp, Z.ci, Acur, H.scratch.p);
    hipLaunchKernelGGL(k_amg_gauss_jordan, dim3(1), dim3(1024), 0, c.stream, (int)n, H.scratch.p,
                       H.scratch.p + 2 * n * n, H.err.p);
    H.cinv.alloc(n * n);
    hipLaunchKernelGGL(k_amg_extract_inv<double>, g1(n * n), dim3(256), 0, c.stream, n, H.scratch.p, H.cinv.p);
    MHX_HIP(hipGetLastError());
    finish_setup(c, H, nlev - 1);
  }
  // block-Jacobi smoother: nodal block inverses of every smoothed level
  if (H.smoother != kSmootherChebyshev) {
    for (int l = 0; l + 1 < nlev; ++l) {
      AmgLevel& V = *H.lv[l];
      const long nblk = V.n / nf;
      V.binv.alloc((size_t)nblk * nf * nf);
      const double* Al = l ? V.A.p : A;
      switch (nf) {
        case 1: hipLaunchKernelGGL(k_amg_block_inv<1>, g1(nblk), dim3(256), 0, c.stream, nblk, V.rp, V.ci, Al, V.binv.p); break;
        case 2: hipLaunchKernelGGL(k_amg_block_inv<2>, g1(nblk), dim3(256), 0, c.stream, nblk, V.rp, V.ci, Al, V.binv.p); break;
        case 3: hipLaunchKernelGGL(k_amg_block_inv<3>, g1(nblk), dim3(256), 0, c.stream, nblk, V.rp, V.ci, Al, V.binv.p); break;
        default: hipLaunchKernelGGL(k_amg_block_inv<4>, g1(nblk), dim3(256), 0, c.stream, nblk, V.rp, V.ci, Al, V.binv.p); break;
      }
    }
  }
  // FP32 copies of the operators
  for (int l = 0; l < nlev; ++l) {
    AmgLevel& V = *H.lv[l];
    const bool last = l == nlev - 1;
    auto vecs = [&](auto& W) {
      W.x.alloc(V.n); W.t.alloc(V.n); W.b.alloc(V.n); W.r.alloc(V.n); W.d.alloc(V.n);
    };
    if (fp32) {
      vecs(V.f32);
      auto cast = [&](const double* s, DevBuf<float>& d, long m) {
        d.alloc(m);
        hipLaunchKernelGGL((k_amg_cast<double, float>), g1(m), dim3(256), 0, c.stream, m, s, d.p);
      };
      cast(V.dinv.p, V.f32.dinv, V.n);
      if (H.smoother != kSmootherChebyshev && !last) cast(V.binv.p, V.f32.binv, V.binv.n);
      if (l > 0) cast(V.A.p, V.f32.A, V.A.n);
      if (!last) {
        cast(V.P.p, V.f32.P, V.P.n);
        cast(V.R.p, V.f32.R, V.R.n);
      }
    } else {
      vecs(V.f64);
    }
  }
  MHX_HIP(hipGetLastError());
}

// Chebyshev coefficients of one level (Ifpack2's three-term recurrence)
struct Cheb {
  double inv_theta, a2, b2;
};
static Cheb cheb_coef(double lambda) {
  const double lmax = kChebBoost * lambda, lmin = lmax / kChebRatio;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta, rho0 = 1.0 / sigma, rho1 = 1.0 / (2.0 * sigma - rho0);
  return Cheb{1.0 / theta, rho1 * rho0, 2.0 * rho1 / delta};
}

template <class T>
struct LvView {
  const T *A, *dinv, *P, *R, *binv;
  T *x, *t, *b, *r, *d;
};

template <class T>
static LvView<T> view(Amg& H, int l, const T* A0);
template <>
LvView<double> view<double>(Amg& H, int l, const double* A0) {
  AmgLevel& V = *H.lv[l];
  return {l ? V.A.p : A0, V.dinv.p, V.P.p, V.R.p, V.binv.p, V.f64.x.p, V.f64.t.p, V.f64.b.p, V.f64.r.p, V.f64.d.p};
}
template <>
LvView<float> view<float>(Amg& H, int l, const float* A0) {
  AmgLevel& V = *H.lv[l];
  return {l ? V.f32.A.p : A0, V.f32.dinv.p, V.f32.P.p, V.f32.R.p, V.f32.binv.p, V.f32.x.p, V.f32.t.p, V.f32.b.p,
          V.f32.r.p, V.f32.d.p};
}

template <class T>
static void cheb_step(Ctx& c, AmgLevel& V, const LvView<T>& W, const T* b, const T* xin, T* xout, double alpha,
                      double beta, int zero_x) {
  by_lanes(zero_x ? 4 : V.lanes, [&](auto LN) {
    by_lat(V.lat, [&](auto LT) {
      constexpr int K = decltype(LN)::value;
      constexpr bool LAT = decltype(LT)::value;
      hipLaunchKernelGGL((k_amg_cheb<K, LAT, T>), g1(V.n * K), dim3(256), 0, c.stream, c.mesh, V.n, V.rp, V.ci, W.A,
                         b, xin, xout, W.d, W.dinv, alpha, beta, zero_x);
    });
  });
}

// r = b - A x on one level
template <class T>
static void level_resid(Ctx& c, AmgLevel& V, const LvView<T>& W, const T* b, const T* x, T* r) {
  by_lanes(V.lanes, [&](auto LN) {
    by_lat(V.lat, [&](auto LT) {
      constexpr int K = decltype(LN)::value;
      constexpr bool LAT = decltype(LT)::value;
      hipLaunchKernelGGL((k_amg_resid<K, LAT, T>), g1(V.n * K), dim3(256), 0, c.stream, c.mesh, V.n, V.rp, V.ci,
                         W.A, b, x, r);
    });
  });
}

// damped nodal block-Jacobi sweeps: x <- x + omega Binv (b - A x); with
// zero_x the first sweep starts from x = 0 and needs no residual
template <class T>
static void bj_smooth(Ctx& c, Amg& H, AmgLevel& V, const LvView<T>& W, const T* b, T* x, int zero_x) {
  for (int s = 0; s < H.sweeps; ++s) {
    const T* r = b;
    if (!(zero_x && s == 0)) {
      level_resid<T>(c, V, W, b, x, W.t);
      r = W.t;
    }
    hipLaunchKernelGGL(k_amg_bj<T>, g1(V.n), dim3(256), 0, c.stream, V.n, H.nf, W.binv, r, x, x, H.omega,
                       (int)(zero_x && s == 0));
  }
}

// one multicolour block Gauss-Seidel sweep over the (reach+1)^d colours of
// the level's node lattice, in lexicographic colour order or its reverse
template <class T>
static void gs_sweep(Ctx& c, Amg& H, AmgLevel& V, const LvView<T>& W, const T* b, T* x, bool forward) {
  const int nf = H.nf;
  const int nfp = nf == 1 ? 1 : nf == 2 ? 2 : 4;
  const int s = std::max(1, V.reach) + 1;
  int nc[3];
  for (int d = 0; d < 3; ++d) nc[d] = V.N[d] > 1 ? s : 1;
  const int ncolour = nc[0] * nc[1] * nc[2];
  const int lanes = std::min(V.lanes, 64 / nfp);
  for (int k = 0; k < ncolour; ++k) {
    const int col = forward ? k : ncolour - 1 - k;
    const int cx = col % nc[0], cy = (col / nc[0]) % nc[1], cz = col / (nc[0] * nc[1]);
    const int co[3] = {cx, cy, cz};
    long cnt[3];
    for (int d = 0; d < 3; ++d) cnt[d] = V.N[d] > co[d] ? (V.N[d] - co[d] + s - 1) / s : 0;
    const long ncol = cnt[0] * cnt[1] * cnt[2];
    if (!ncol) continue;
    by_lanes(lanes, [&](auto LN) {
      constexpr int K = decltype(LN)::value;
      auto go = [&](auto NP) {
        constexpr int F = decltype(NP)::value;
        if constexpr (K * F <= 64)
          hipLaunchKernelGGL((k_amg_gs<K, F, T>), g1(ncol * K * F), dim3(256), 0, c.stream, ncol, nf, s, cx, cy, cz,
                             cnt[0], cnt[1], (long)V.N[0], (long)V.N[1], V.rp, V.ci, W.A, W.binv, b, x);
        else
          throw Error{MHX_ERR_STATE, "AMG Gauss-Seidel: node group wider than a wavefront"};
      };
      if (nfp == 1) go(std::integral_constant<int, 1>{});
      else if (nfp == 2) go(std::integral_constant<int, 2>{});
      else go(std::integral_constant<int, 4>{});
    });
  }
}

template <class T>
static void vcycle(Ctx& c, Amg& H, const T* A0, const double* cinv, const T* v, T* out) {
  const int nl = (int)H.lv.size();
  // downward: pre-smooth (zero start), residual, restrict
  for (int l = 0; l < nl - 1; ++l) {
    AmgLevel& V = *H.lv[l];
    const LvView<T> W = view<T>(H, l, A0);
    const T* b = l ? W.b : v;
    T* x = l ? W.x : out;
    if (H.smoother == kSmootherBlockJacobi) {
      bj_smooth<T>(c, H, V, W, b, x, 1);
    } else if (H.smoother == kSmootherBlockGaussSeidel) {
      hipLaunchKernelGGL(k_amg_fill0<T>, g1(V.n), dim3(256), 0, c.stream, V.n, x);
      gs_sweep<T>(c, H, V, W, b, x, true);
    } else {
      const Cheb k = cheb_coef(V.lambda);
      cheb_step<T>(c, V, W, b, x, W.t, 0.0, k.inv_theta, 1);
      cheb_step<T>(c, V, W, b, W.t, x, k.a2, k.b2, 0);
    }
    level_resid<T>(c, V, W, b, x, W.r);
    const LvView<T> Wc = view<T>(H, l + 1, A0);
    const long nc = H.lv[l + 1]->n;
    by_lanes(V.rlane, [&](auto LN) {
      constexpr int K = decltype(LN)::value;
      hipLaunchKernelGGL((k_amg_spmv<K, T>), g1(nc * K), dim3(256), 0, c.stream, nc, V.rrp.p, V.rci.p, W.R, W.r,
                         Wc.b, 0);
    });
  }
  // coarsest
  {
    AmgLevel& Z = *H.lv[nl - 1];
    const LvView<T> W = view<T>(H, nl - 1, A0);
    const T* b = nl > 1 ? W.b : v;
    T* x = nl > 1 ? W.x : out;
    hipLaunchKernelGGL(k_amg_gemv<T>, g1(Z.n * 64), dim3(256), 0, c.stream, Z.n, cinv, b, x);
  }
  // upward: prolong-add, post-smooth
  for (int l = nl - 2; l >= 0; --l) {
    AmgLevel& V = *H.lv[l];
    const LvView<T> W = view<T>(H, l, A0);
    const T* b = l ? W.b : v;
    T* x = l ? W.x : out;
    const LvView<T> Wc = view<T>(H, l + 1, A0);
    by_lanes(V.plane, [&](auto LN) {
      constexpr int K = decltype(LN)::value;
      hipLaunchKernelGGL((k_amg_spmv<K, T>), g1(V.n * K), dim3(256), 0, c.stream, V.n, V.prp.p, V.pci.p, W.P, Wc.x,
                         x, 1);
    });
    if (H.smoother == kSmootherBlockJacobi) {
      bj_smooth<T>(c, H, V, W, b, x, 0);
    } else if (H.smoother == kSmootherBlockGaussSeidel) {
      gs_sweep<T>(c, H, V, W, b, x, false);
    } else {
      const Cheb k = cheb_coef(V.lambda);
      cheb_step<T>(c, V, W, b, x, W.t, 0.0, k.inv_theta, 0);
      cheb_step<T>(c, V, W, b, W.t, x, k.a2, k.b2, 0);
    }
  }
  MHX_HIP(hipGetLastError());
}

void amg_apply(Ctx& c, const double* A, const double* v, double* out) {
  MHX_REQUIRE(c.amg && !c.amg->lv.empty() && !c.amg->fp32, MHX_ERR_STATE, "AMG: no FP64 hierarchy");
  ScopedPhase ph(c.prof, "MueLu::Hierarchy::Iterate", c.stream);
  vcycle<double>(c, *c.amg, A, c.amg->cinv.p, v, out);
}
void amg_apply(Ctx& c, const float* A, const float* v, float* out) {
  MHX_REQUIRE(c.amg && !c.amg->lv.empty() && c.amg->fp32, MHX_ERR_STATE, "AMG: no FP32 hierarchy");
  ScopedPhase ph(c.prof, "MueLu::Hierarchy::Iterate", c.stream);
  vcycle<float>(c, *c.amg, A, c.amg->cinv.p, v, out);
}

int amg_levels(const Ctx& c) { return c.amg ? (int)c.amg->lv.size() : 0; }

// sizes (rows, cols, nnz) of one operator of the hierarchy; copies it to the
// host arrays that are not null.  which: 0 = A_l, 1 = P_l, 2 = R_l, 3 = the
// coarsest dense inverse (row-major, rows = cols = its size).
void amg_download(Ctx& c, int level, int which, int64_t* rows, int64_t* cols, int64_t* nnz, int32_t* rp, int32_t* ci,
                  double* vals) {
  MHX_REQUIRE(c.amg && level >= 0 && level < (int)c.amg->lv.size(), MHX_ERR_ARG, "AMG: no such level");
  Amg& H = *c.amg;
  AmgLevel& V = *H.lv[level];
  const bool last = level == (int)H.lv.size() - 1;
  const int32_t *prp = nullptr, *pci = nullptr;
  const double* pv = nullptr;
  long nr = 0, nc = 0, nz = 0;
  if (which == 0) {
    prp = V.rp; pci = V.ci; pv = level ? V.A.p : H.A0;
    nr = nc = V.n;
  } else if (which == 1 || which == 2) {
    MHX_REQUIRE(!last, MHX_ERR_ARG, "AMG: the coarsest level has no transfer operators");
    const long m = H.lv[level + 1]->n;
    if (which == 1) { prp = V.prp.p; pci = V.pci.p; pv = V.P.p; nr = V.n; nc = m; }
    else { prp = V.rrp.p; pci = V.rci.p; pv = V.R.p; nr = m; nc = V.n; }
  } else if (which == 3) {
    MHX_REQUIRE(last, MHX_ERR_ARG, "AMG: the dense inverse belongs to the coarsest level");
    nr = nc = V.n;
    nz = V.n * V.n;
  } else {
    throw Error{MHX_ERR_ARG, "AMG: unknown operator"};
  }
  if (which != 3) {
    int32_t last_rp = 0;
    MHX_HIP(hipMemcpyAsync(&last_rp, prp + nr, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    MHX_HIP(hipStreamSynchronize(c.stream));
    nz = last_rp;
    if (rp) MHX_HIP(hipMemcpyAsync(rp, prp, (nr + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    if (ci) MHX_HIP(hipMemcpyAsync(ci, pci, nz * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    if (vals) MHX_HIP(hipMemcpyAsync(vals, pv, nz * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  } else if (vals) {
    MHX_HIP(hipMemcpyAsync(vals, H.cinv.p, nz * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  }
  MHX_HIP(hipStreamSynchronize(c.stream));
  if (rows) *rows = nr;
  if (cols) *cols = nc;
  if (nnz) *nnz = nz;
}

double amg_lambda(const Ctx& c, int level) {
  return (c.amg && level >= 0 && level < (int)c.amg->lv.size()) ? c.amg->lv[level]->lambda : 0.0;
}

}  // namespace mhx
