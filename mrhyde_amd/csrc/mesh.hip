// mesh.hip — structured box mesh, DOF numbering, reference 1-D rules,
// strong-Dirichlet tables and the CRS graph (row pointer on the host, column
// indices by a device kernel).
//
// Replaces the setup half of the reference path:
//   MeshInterface inline Square/Cube factories (src/interfaces/mesh/meshInterface_construct.hpp:134-200)
//   DiscretizationInterface::buildDOFManagers (discretizationInterface_dof.hpp:14-241)
//   getBasis / getQuadrature (discretizationInterface_basis.hpp:219-350, _integration.hpp:91-106)
//   setDirichletData (discretizationInterface_dof.hpp:643-736)
//   setupLinearAlgebra graph (linearAlgebraInterface_construct.hpp:140-265)
//
// Numbering: DOF nodes form a (p*NX+1) x (p*NY+1) [x (p*NZ+1)] lattice,
// numbered lexicographically (x fastest).  Element e = ex + NX*(ey + NY*ez)
// holds lattice nodes p*e_d + {0..p}; its local order is lexicographic too.
// The CRS row of node r has the columns of every element sharing r; on the
// lattice that is a tensor product of 1-D ranges [lo_d, hi_d], so sorted
// column order is z-major, then y, then x.

#include <cmath>
#include <cstring>

#include "mhx_internal.hpp"

namespace mhx {

// Gauss-Legendre on [-1,1], points in descending order.
void gauss_legendre_desc(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double p0 = 1.0, p1 = 0.0, dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      p0 = 1.0;
      p1 = 0.0;
      for (int k = 1; k <= n; ++k) {
        double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
      }
      dp = n * (z * p0 - p1) / (z * z - 1.0);
      double dz = p0 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    p0 = 1.0;
    p1 = 0.0;
    for (int k = 1; k <= n; ++k) {
      double p2 = p1;
      p1 = p0;
      p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
    }
    dp = n * (z * p0 - p1) / (z * z - 1.0);
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

// Equispaced Lagrange polynomials of order p on [-1,1] and their slopes.
void lagrange_equispaced(int p, double xi, double* L, double* dL) {
  double nodes[kMaxP1];
  for (int k = 0; k <= p; ++k) nodes[k] = -1.0 + 2.0 * k / p;
  for (int k = 0; k <= p; ++k) {
    double v = 1.0, d = 0.0;
    for (int m = 0; m <= p; ++m) {
      if (m == k) continue;
      double den = nodes[k] - nodes[m];
      double prod = 1.0 / den;
      for (int l = 0; l <= p; ++l) {
        if (l == k || l == m) continue;
        prod *= (xi - nodes[l]) / (nodes[k] - nodes[l]);
      }
      d += prod;
      v *= (xi - nodes[m]) / den;
    }
    L[k] = v;
    dL[k] = d;
  }
}

static inline double u01(uint64_t seed, uint64_t idx) {
  uint64_t z = seed ^ (idx * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// One thread per DOF row: the row's columns are the nf unknowns of every
// lattice node in the tensor range [lo_d, hi_d], node-major, field-minor.
__global__ void k_fill_colidx(BoxMesh m, const int32_t* __restrict__ rowptr, int32_t* __restrict__ colidx) {
  long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m.ndof) return;
  const long node = r / m.nf;
  int c[3] = {(int)(node % m.nn[0]), (int)((node / m.nn[0]) % m.nn[1]), (int)(node / ((long)m.nn[0] * m.nn[1]))};
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    if (a >= m.dim) { lo[a] = hi[a] = 0; continue; }
    if (c[a] % m.p == 0) {
      lo[a] = c[a] - m.p < 0 ? 0 : c[a] - m.p;
      hi[a] = c[a] + m.p > m.nn[a] - 1 ? m.nn[a] - 1 : c[a] + m.p;
    } else {
      lo[a] = (c[a] / m.p) * m.p;
      hi[a] = lo[a] + m.p;
    }
  }
  long k = rowptr[r];
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const long nb = x + (long)m.nn[0] * (y + (long)m.nn[1] * z);
        for (int f = 0; f < m.nf; ++f) colidx[k++] = (int32_t)(nb * m.nf + f);
      }
}

static inline long row_len_1d(int c, int p, int n) {
  if (c % p == 0) {
    int lo = c - p < 0 ? 0 : c - p;
    int hi = c + p > n - 1 ? n - 1 : c + p;
    return hi - lo + 1;
  }
  return p + 1;
}

void build_box(Ctx& c) {
  const mhx_problem_desc& d = c.desc;
  BoxMesh& m = c.mesh;
  m.dim = d.dim;
  m.p = d.order;
  m.nq1 = d.quadrature / 2 + 1;
  MHX_REQUIRE(m.p + 1 <= kMaxP1 && m.nq1 <= kMaxQ1, MHX_ERR_UNSUPPORTED, "order/quadrature too high");
  m.nel[0] = d.nx;
  m.nel[1] = d.ny;
  m.nel[2] = d.dim == 3 ? d.nz : 1;
  const bool ns = d.physics == MHX_PHYS_NAVIER_STOKES;
  m.nf = ns ? m.dim + 1 : 1;
  c.nfield = m.nf;
  m.ne = m.nnode = m.nvert = 1;
  for (int a = 0; a < 3; ++a) {
    bool act = a < m.dim;
    m.nn[a] = act ? m.p * m.nel[a] + 1 : 1;
    m.nvx[a] = act ? m.nel[a] + 1 : 1;
    m.ne *= m.nel[a];
    m.nnode *= m.nn[a];
    m.nvert *= m.nvx[a];
  }
  m.ndof = m.nnode * m.nf;
  MHX_REQUIRE(m.ndof < (1L << 31) - 1, MHX_ERR_UNSUPPORTED, "more than 2^31 DOFs");
  c.nb = 1;
  c.nq = 1;
  for (int a = 0; a < m.dim; ++a) {
    c.nb *= (m.p + 1);
    c.nq *= m.nq1;
  }
  // 1-D tables
  std::memset(&c.tab, 0, sizeof(c.tab));
  double xq[kMaxQ1], wq[kMaxQ1];
  gauss_legendre_desc(m.nq1, xq, wq);
  for (int q = 0; q < m.nq1; ++q) {
    double L[kMaxP1], dL[kMaxP1], N[2], dN[2];
    lagrange_equispaced(m.p, xq[q], L, dL);
    lagrange_equispaced(1, xq[q], N, dN);
    for (int i = 0; i <= m.p; ++i) {
      c.tab.B[i][q] = L[i];
      c.tab.D[i][q] = dL[i];
    }
    for (int i = 0; i < 2; ++i) {
      c.tab.N[i][q] = N[i];
      c.tab.M[i][q] = dN[i];
    }
    c.tab.W[q] = wq[q];
  }
  // vertices (optionally jittered in the interior)
  double lo[3] = {d.xmin, d.ymin, d.zmin}, hi[3] = {d.xmax, d.ymax, d.zmax};
  c.h_vx.assign(m.nvert * 3, 0.0);
  for (long v = 0; v < m.nvert; ++v) {
    long vc[3] = {v % m.nvx[0], (v / m.nvx[0]) % m.nvx[1], v / ((long)m.nvx[0] * m.nvx[1])};
    bool interior = true;
    for (int a = 0; a < m.dim; ++a)
      if (vc[a] == 0 || vc[a] == m.nvx[a] - 1) interior = false;
    for (int a = 0; a < m.dim; ++a) {
      double h = (hi[a] - lo[a]) / m.nel[a];
      double x = lo[a] + vc[a] * h;
      if (interior && d.perturb != 0.0) x += d.perturb * h * (2.0 * u01(d.perturb_seed, (uint64_t)v * 3 + a) - 1.0);
      c.h_vx[v * 3 + a] = x;
    }
  }
  // Dirichlet mask and values.  Thermal: the DOFs on dbc_sides.  Navier-
  // Stokes: velocity component c on dbc_sides | ns_lid_sides |
  // ns_comp_sides[c] (ns_lid_u on a lid side, edges included, else 0) and
  // optionally the pressure at node 0.
  c.h_fixed.assign(m.ndof, 0);
  c.h_dbc.assign(m.ndof, 0.0);
  c.h_dbc_list.clear();
  const uint32_t vel_sides = d.dbc_sides | (ns ? d.ns_lid_sides : 0u);
  for (long node = 0; node < m.nnode; ++node) {
    long gc[3] = {node % m.nn[0], (node / m.nn[0]) % m.nn[1], node / ((long)m.nn[0] * m.nn[1])};
    uint32_t on = 0;
    for (int a = 0; a < m.dim; ++a) {
      if (gc[a] == 0) on |= 1u << (2 * a);
      if (gc[a] == m.nn[a] - 1) on |= 1u << (2 * a + 1);
    }
    for (int f = 0; f < m.nf; ++f) {
      const long g = node * m.nf + f;
      bool fix = false;
      double val = 0.0;
      if (!ns) {
        fix = (on & d.dbc_sides) != 0;
        val = d.dbc_values ? d.dbc_values[g] : d.dbc_const;
      } else if (f < m.dim) {
        fix = (on & (vel_sides | d.ns_comp_sides[f])) != 0;
        val = (on & d.ns_lid_sides) ? d.ns_lid_u[f] : 0.0;
      } else {
        fix = d.ns_pin_pressure && node == 0;
      }
      if (fix) {
        c.h_fixed[g] = 1;
        c.h_dbc[g] = val;
        c.h_dbc_list.push_back((int32_t)g);
      }
    }
  }
  // CRS row pointer
  c.h_rowptr.assign(m.ndof + 1, 0);
  long acc = 0;
  for (long g = 0; g < m.ndof; ++g) {
    const long node = g / m.nf;
    long gc[3] = {node % m.nn[0], (node / m.nn[0]) % m.nn[1], node / ((long)m.nn[0] * m.nn[1])};
    long len = m.nf;
    for (int a = 0; a < m.dim; ++a) len *= row_len_1d((int)gc[a], m.p, m.nn[a]);
    c.h_rowptr[g] = (int32_t)acc;
    acc += len;
    MHX_REQUIRE(acc < (1L << 31) - 1, MHX_ERR_UNSUPPORTED, "more than 2^31 nonzeros");
  }
  c.h_rowptr[m.ndof] = (int32_t)acc;
  c.nnz = acc;

  // device copies
  c.rowptr.alloc(m.ndof + 1);
  c.colidx.alloc(c.nnz);
  c.fixed.alloc(m.ndof);
  c.vx.alloc(m.nvert * 3);
  c.dbc_vals.alloc(m.ndof);
  c.dbc_list.alloc(c.h_dbc_list.size() ? c.h_dbc_list.size() : 1);
  MHX_HIP(hipMemcpyAsync(c.rowptr.p, c.h_rowptr.data(), (m.ndof + 1) * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
  MHX_HIP(hipMemcpyAsync(c.fixed.p, c.h_fixed.data(), m.ndof, hipMemcpyHostToDevice, c.stream));
  MHX_HIP(hipMemcpyAsync(c.vx.p, c.h_vx.data(), m.nvert * 3 * sizeof(double), hipMemcpyHostToDevice, c.stream));
  MHX_HIP(hipMemcpyAsync(c.dbc_vals.p, c.h_dbc.data(), m.ndof * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if (!c.h_dbc_list.empty())
    MHX_HIP(hipMemcpyAsync(c.dbc_list.p, c.h_dbc_list.data(), c.h_dbc_list.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, c.stream));
  if (d.k0) {
    c.k0.alloc(m.ne);
    MHX_HIP(hipMemcpyAsync(c.k0.p, d.k0, m.ne * sizeof(double), hipMemcpyHostToDevice, c.stream));
  }
  if (d.src_elem) {
    c.src.alloc(m.ne);
    MHX_HIP(hipMemcpyAsync(c.src.p, d.src_elem, m.ne * sizeof(double), hipMemcpyHostToDevice, c.stream));
  }
  long blocks = (m.ndof + 255) / 256;
  hipLaunchKernelGGL(k_fill_colidx, dim3((unsigned)blocks), dim3(256), 0, c.stream, m, c.rowptr.p, c.colidx.p);
  MHX_HIP(hipGetLastError());
  MHX_HIP(hipStreamSynchronize(c.stream));
  // pattern width for SpMV: a power of two in [4,64] giving each lane about
  // 8 to 14 entries (two or three four-entry passes).  Measured on MI355X:
  // mean row 27 (Q1 hex) -> 4 lanes, 63 (Q2 hex) -> 8, 105 (Q1 hex, 4 fields) -> 8.
  double avg = (double)c.nnz / (double)m.ndof;
  int lanes = 4;
  while (lanes < 64 && lanes * 14 < avg) lanes *= 2;
  c.spmv_lanes = lanes;
}

}  // namespace mhx
