// lattice.hpp — element / node arithmetic on the structured box lattice and
// the colour schedule shared by the assembly kernels.
#pragma once

#include "mhx_internal.hpp"

namespace mhx {

// lattice helpers ------------------------------------------------------------
template <int DIM, int P>
__device__ __forceinline__ long node_of(const BoxMesh& m, const int* ec, int j) {
  constexpr int P1 = P + 1;
  int jx = j % P1, jy = (j / P1) % P1, jz = DIM == 3 ? j / (P1 * P1) : 0;
  return (long)(P * ec[0] + jx) + (long)m.nn[0] * ((long)(P * ec[1] + jy) + (long)m.nn[1] * (long)(P * ec[2] + jz));
}

// position of column-node offset (local jl) inside the row of local node il,
// along one axis: returns pos and the 1-D row length.
template <int P>
__device__ __forceinline__ void axis_slot(int il, int jl, int ecd, int nnd, int& pos, int& len) {
  if (il == 0 || il == P) {
    int r = P * ecd + il;
    int lo = r - P < 0 ? 0 : r - P;
    int hi = r + P > nnd - 1 ? nnd - 1 : r + P;
    pos = P * ecd + jl - lo;
    len = hi - lo + 1;
  } else {
    pos = jl;
    len = P + 1;
  }
}

// Element owned by group g of this workgroup in sub-phase `sub`.
//   element colouring (a.blocked == 0): one launch per colour, one element per
//     group, elements of the colour enumerated over the grid;
//   block colouring (a.blocked == 1): the mesh is tiled by blocks of
//     bw = 2*s elements per axis; one launch per block colour, one block per
//     workgroup; sub-phase `sub` (an element colour) gives group g the element
//     block_origin + 2*g_d + sub_d.  Same-colour elements of one block share no
//     DOF, and same-colour blocks are a block apart, so no two concurrent
//     element scatters touch the same CRS entry.
template <int DIM>
__device__ __forceinline__ bool element_of(const AsmArgs& a, int g, int sub, int* ec) {
  ec[0] = ec[1] = ec[2] = 0;
  if (!a.blocked) {
    const int idx = (int)blockIdx.x * a.gpb + g;
    int t = idx < a.ncolor ? idx : 0;
    ec[0] = 2 * (t % a.cn[0]) + (a.color & 1);
    t /= a.cn[0];
    ec[1] = 2 * (t % a.cn[1]) + ((a.color >> 1) & 1);
    t /= a.cn[1];
    if (DIM == 3) ec[2] = 2 * t + ((a.color >> 2) & 1);
    return idx < a.ncolor;
  }
  int t = blockIdx.x;
  const int gc[3] = {g % a.s[0], (g / a.s[0]) % a.s[1], g / (a.s[0] * a.s[1])};
  bool ok = true;
  for (int d = 0; d < DIM; ++d) {
    const int bi = 2 * (t % a.cn[d]) + ((a.color >> d) & 1);
    t /= a.cn[d];
    ec[d] = bi * a.bw[d] + 2 * gc[d] + ((sub >> d) & 1);
    ok = ok && ec[d] < a.m.nel[d];
  }
  return ok && g < a.gpb;
}

// Segment scheduling (a.blocked == 2): a workgroup owns a segment of a.gpb
// x-consecutive elements at one (ey, ez); group g takes the segment's g-th
// element.  One launch per colour of (segment parity, ey parity, ez parity):
// same-colour segments are a segment apart in x or an element apart in y/z,
// so they share no node.  Returns whether group g has an element (the last
// segment of a row may be short) and the segment index and length.
template <int DIM>
__device__ __forceinline__ bool element_of_seg(const AsmArgs& a, int g, int* ec, int& sx, int& s_act) {
  int t = blockIdx.x;
  sx = 2 * (t % a.cn[0]) + (a.color & 1);
  t /= a.cn[0];
  ec[1] = 2 * (t % a.cn[1]) + ((a.color >> 1) & 1);
  t /= a.cn[1];
  ec[2] = DIM == 3 ? 2 * t + ((a.color >> 2) & 1) : 0;
  ec[0] = sx * a.gpb + g;
  const int rest = a.m.nel[0] - sx * a.gpb;
  s_act = rest < a.gpb ? rest : a.gpb;
  return g < s_act;
}

// Scheduling key of element index e along one axis: (block colour bit,
// element colour bit).  Launch order is block colour major, sub-phase minor,
// so along each axis the smaller key runs first.
__device__ __forceinline__ int axis_key(const AsmArgs& a, int d, int e) {
  return (a.blocked == 1 ? ((e / a.bw[d]) & 1) * 2 : 0) + (e & 1);
}

// Columns of a row of the context's lattice pattern in closed form (the fill
// order of k_fill_colidx: a box of nodes, z-major, fields innermost), so the
// sparse products can skip the 4-byte column index stream.  Entry q of the
// row: run = q / (ex nf) selects the (y, z) line of the box, the rest the node
// along x and the field.  The divisions by these small extents use float
// reciprocals: for q < 2^12 the product is within 1e-3 of (q + 0.5) / w.
struct LatRow {
  int lo0, lo1, lo2, ex, ey, nf, nn0, nn1;
  float iw, iey, inf;
  __device__ __forceinline__ int col(int q) const {
    const int w = ex * nf;
    const int run = (int)(((float)q + 0.5f) * iw);
    const int r0 = q - run * w;
    const int iz = (int)(((float)run + 0.5f) * iey);
    const int iy = run - iz * ey;
    const int dx = nf == 1 ? r0 : (int)(((float)r0 + 0.5f) * inf);
    const int f = r0 - dx * nf;
    return (((lo2 + iz) * nn1 + (lo1 + iy)) * nn0 + (lo0 + dx)) * nf + f;
  }
};

__device__ __forceinline__ LatRow lat_row(const BoxMesh& m, long row) {
  // row and node numbers fit 32 bits (int32 CRS); 32-bit unsigned division
  // is several times cheaper than the 64-bit one
  const unsigned node = (unsigned)row / (unsigned)m.nf;
  const unsigned yz = node / (unsigned)m.nn[0];
  const unsigned zz = yz / (unsigned)m.nn[1];
  const int c[3] = {(int)(node - yz * (unsigned)m.nn[0]), (int)(yz - zz * (unsigned)m.nn[1]), (int)zz};
  int lo[3], ext[3];
  for (int a = 0; a < 3; ++a) {
    if (a >= m.dim) {
      lo[a] = 0;
      ext[a] = 1;
    } else if (c[a] % m.p == 0) {
      lo[a] = c[a] - m.p < 0 ? 0 : c[a] - m.p;
      const int hi = c[a] + m.p > m.nn[a] - 1 ? m.nn[a] - 1 : c[a] + m.p;
      ext[a] = hi - lo[a] + 1;
    } else {
      lo[a] = (c[a] / m.p) * m.p;
      ext[a] = m.p + 1;
    }
  }
  LatRow r;
  r.lo0 = lo[0];
  r.lo1 = lo[1];
  r.lo2 = lo[2];
  r.ex = ext[0];
  r.ey = ext[1];
  r.nf = m.nf;
  r.nn0 = m.nn[0];
  r.nn1 = m.dim > 1 ? m.nn[1] : 1;
  r.iw = 1.0f / (float)(ext[0] * m.nf);
  r.iey = 1.0f / (float)ext[1];
  r.inf = 1.0f / (float)m.nf;
  return r;
}

// column source of a sparse product: the stored index stream, or the closed
// form of the context's lattice pattern
struct CsrCols {
  const int32_t* ci;
  __device__ __forceinline__ int operator()(int k) const { return ci[k]; }
};
struct LatCols {
  LatRow r;
  int b;  // first entry of the row
  __device__ __forceinline__ int operator()(int k) const { return r.col(k - b); }
};

}  // namespace mhx
