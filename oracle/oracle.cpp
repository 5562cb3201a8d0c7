// oracle.cpp — CPU restatement of MrHyDE's element-assembly + Newton/Krylov path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: it is
// imported by tests/, by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg, and nowhere else.  The product library (mrhyde_amd/)
// never links, loads or calls it.
//
// It restates the reference algorithm in plain scalar C++ (FP64), following
// the reference files cited at each function, with a Sacado-SFad-style dual
// number.  Pinning (what checks each part against something other than a
// retyped formula) is exercised by tests/test_oracle.py:
//   basis/quadrature/geometry  -> regression/discretization/HGRAD/mrhyde.gold
//   thermal weak form, DBCs    -> regression/thermal/*/mrhyde.gold L2 errors
//   dual-number Jacobian       -> central finite differences of the residual
//   CRS graph                  -> closed-form nnz (3n+1)^d / (8n+1)^d
//   Krylov                     -> scipy.sparse.linalg.spsolve on the same system
//
// The local DOF order inside an element is tensor-lexicographic (x fastest);
// the reference's Q1 vertex order is a fixed permutation of it (applied by the
// KAT test).  Global DOFs are numbered lexicographically on the node lattice.

#include "../include/mhx.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// SFad<N>: value + N derivative slots, forward-mode product/chain rules
// (Sacado::Fad::SFad semantics; typedefs src/preferences.hpp:79-100,
//  seeding EvalT(size, index, value) src/tools/workset.cpp:884).
// ---------------------------------------------------------------------------
template <int N>
struct SFad {
  double v;
  double d[N];
  SFad() : v(0.0) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
  SFad(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
  SFad(int, int idx, double x) : v(x) {
    for (int i = 0; i < N; ++i) d[i] = 0.0;
    d[idx] = 1.0;
  }
  SFad& operator+=(const SFad& b) { v += b.v; for (int i = 0; i < N; ++i) d[i] += b.d[i]; return *this; }
  SFad& operator+=(double b) { v += b; return *this; }
};
template <int N> inline SFad<N> operator+(const SFad<N>& a, const SFad<N>& b) { SFad<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> inline SFad<N> operator-(const SFad<N>& a, const SFad<N>& b) { SFad<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> inline SFad<N> operator*(const SFad<N>& a, const SFad<N>& b) { SFad<N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> inline SFad<N> operator/(const SFad<N>& a, const SFad<N>& b) { SFad<N> r; r.v = a.v / b.v; for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v); return r; }
template <int N> inline SFad<N> operator+(const SFad<N>& a, double b) { SFad<N> r = a; r.v += b; return r; }
template <int N> inline SFad<N> operator+(double a, const SFad<N>& b) { SFad<N> r = b; r.v += a; return r; }
template <int N> inline SFad<N> operator-(const SFad<N>& a, double b) { SFad<N> r = a; r.v -= b; return r; }
template <int N> inline SFad<N> operator-(double a, const SFad<N>& b) { SFad<N> r; r.v = a - b.v; for (int i = 0; i < N; ++i) r.d[i] = -b.d[i]; return r; }
template <int N> inline SFad<N> operator*(const SFad<N>& a, double b) { SFad<N> r; r.v = a.v * b; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b; return r; }
template <int N> inline SFad<N> operator*(double a, const SFad<N>& b) { SFad<N> r; r.v = a * b.v; for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i]; return r; }
template <int N> inline SFad<N> operator-(const SFad<N>& a) { SFad<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> inline SFad<N> operator/(const SFad<N>& a, double b) { SFad<N> r; r.v = a.v / b; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b; return r; }
template <int N> inline SFad<N> operator/(double a, const SFad<N>& b) { SFad<N> r; r.v = a / b.v; for (int i = 0; i < N; ++i) r.d[i] = -a * b.d[i] / (b.v * b.v); return r; }
template <int N> inline SFad<N> sqrt(const SFad<N>& a) { SFad<N> r; r.v = std::sqrt(a.v); double s = 0.5 / r.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s; return r; }
inline double dsqrt(double a) { return std::sqrt(a); }
template <int N> inline SFad<N> dsqrt(const SFad<N>& a) { return sqrt(a); }

template <int N> inline double val(const SFad<N>& a) { return a.v; }
template <int N> inline double dx(const SFad<N>& a, int i) { return a.d[i]; }
inline double val(double a) { return a; }
inline double dx(double, int) { return 0.0; }
inline double seeded(double x, int, int) { return x; }

template <class T> struct Seed;
template <int N> struct Seed<SFad<N>> { static SFad<N> make(int idx, double x) { return SFad<N>(N, idx, x); } };
template <> struct Seed<double> { static double make(int, double x) { return x; } };

}  // namespace orc

#include "functions.hpp"

namespace orc {

// ---------------------------------------------------------------------------
// 1-D rules
// ---------------------------------------------------------------------------
// Gauss-Legendre points on [-1,1] in DESCENDING order (the order that
// reproduces regression/discretization/HGRAD/mrhyde.gold), with weights.
// n = degree/2 + 1 points for a cubature of the given degree
// (DefaultCubatureFactory, discretizationInterface_integration.hpp:91-106).
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = 0.0;
      for (int k = 1; k <= n; ++k) {
        double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
      }
      double dp = n * (z * p0 - p1) / (z * z - 1.0);
      double dz = p0 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) {
        // one more evaluation for the weight
      }
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = 0.0;
    for (int k = 1; k <= n; ++k) {
      double p2 = p1;
      p1 = p0;
      p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
    }
    double dp = n * (z * p0 - p1) / (z * z - 1.0);
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

// Equispaced Lagrange basis of order p on [-1,1] (HGRAD_*_Cn, EQUISPACED,
// discretizationInterface_basis.hpp:219-350): value and derivative of L_k at xi.
static void lagrange1d(int p, double xi, double* L, double* dL) {
  std::vector<double> nodes(p + 1);
  for (int k = 0; k <= p; ++k) nodes[k] = -1.0 + 2.0 * k / p;
  for (int k = 0; k <= p; ++k) {
    double val = 1.0, der = 0.0;
    for (int m = 0; m <= p; ++m) {
      if (m == k) continue;
      double den = nodes[k] - nodes[m];
      double prod = 1.0 / den;
      for (int l = 0; l <= p; ++l) {
        if (l == k || l == m) continue;
        prod *= (xi - nodes[l]) / (nodes[k] - nodes[l]);
      }
      der += prod;
      val *= (xi - nodes[m]) / den;
    }
    L[k] = val;
    dL[k] = der;
  }
}

static inline double rng_u01(uint64_t seed, uint64_t idx) {
  uint64_t z = seed ^ (idx * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// Problem: structured box mesh, DOFs, reference tables, CRS graph, DBCs.
// ---------------------------------------------------------------------------
struct Problem {
  mhx_problem_desc desc;
  int dim, p, nq1, nb, nq, nv;       // nb = basis functions/elem, nq = points/elem, nv = vertices/elem
  int nf;                            // unknowns per node (thermal 1; Navier-Stokes dim+1)
  int nel[3], nn[3], nvx[3];         // elements, DOF nodes, vertices per axis
  long ne, ndof, nvert, nnode;       // ndof = nnode * nf, DOF = node * nf + field
  std::vector<double> k0, src;       // per element
  int threads = 1;                   // OpenMP threads of the thermal assembly (1: sequential order)
  FManager fm;                       // coefficient functions (FunctionManager)
  bool fm_active = false;            // a thermal coefficient is given as a function
  std::vector<double> vx;            // vertex coords [nvert*3]
  std::vector<int> lids;             // [ne*nf*nb], (e, field, basis)
  std::vector<double> qx1, qw1;      // 1-D Gauss points/weights
  std::vector<double> bref, gref;    // [nb*nq], [nb*nq*dim] reference basis / grads
  std::vector<double> qwref;         // [nq]
  std::vector<double> vbas, vgrad;   // Q1 geometry map [nv*nq], [nv*nq*dim]
  std::vector<int> rowptr, colidx;   // CRS
  std::vector<char> fixed;           // strong-DBC mask
  std::vector<double> dbcval;        // [ndof]
  int ad_size;
  // transient state (Workset::computeSolnTransientSeeded, workset.cpp:601-680):
  // per DOF the evaluated solution alpha_u u + beta_u and u_dot = alpha_t u +
  // beta_t, u the stage unknown (one-stage tableaus: alpha_u = 1, no beta_u)
  int bdf = 0;
  double alpha_t = 0.0;
  std::vector<double> beta_t;        // [ndof] (sum_k BDF_k u_prev[k-1]) / (dt b_s)
  double alpha_u = 1.0;              // A_ss / b_s
  std::vector<double> beta_u;        // [ndof] or empty

  void vertex_of_elem(long e, int v, double* X) const;
  void elem_coords(long e, long* c) const {
    c[0] = e % nel[0];
    c[1] = (e / nel[0]) % nel[1];
    c[2] = dim == 3 ? e / ((long)nel[0] * nel[1]) : 0;
  }
};

static std::string g_err;

// the deck's coefficient functions (physicsInterface_functions.hpp:94-150,
// thermal.cpp:55-61): every entry registered, decomposed at once
static const char* const kThermalFunctions[] = {"thermal diffusion", "thermal source", "specific heat", "density",
                                                "advection x", "advection y", "advection z"};
static void set_functions(Problem& P, const mhx_problem_desc& d) {
  P.fm = FManager();
  P.fm_active = false;
  if (d.n_functions <= 0) return;
  if (!d.function_names || !d.function_exprs) ffail("function names/expressions missing");
  P.fm.have_k0 = d.k0 != nullptr;
  P.fm.have_src = d.src_elem != nullptr;
  for (int i = 0; i < d.n_functions; ++i) P.fm.add(d.function_names[i], d.function_exprs[i]);
  P.fm.decompose();
  for (const char* nm : kThermalFunctions)
    if (P.fm.find(nm) >= 0) P.fm_active = true;
}

// Reference-element tables and mesh (buildDOFManagers, getBasis,
// getQuadrature, SimpleMeshManager-style structured numbering).
static Problem* build_problem(const mhx_problem_desc* d) {
  Problem* P = new Problem();
  P->desc = *d;
  P->dim = d->dim;
  P->p = d->order;
  P->nq1 = d->quadrature / 2 + 1;
  int dim = P->dim, p = P->p;
  P->nb = 1; P->nq = 1; P->nv = 1;
  for (int a = 0; a < dim; ++a) { P->nb *= (p + 1); P->nq *= P->nq1; P->nv *= 2; }
  P->nel[0] = d->nx; P->nel[1] = d->ny; P->nel[2] = dim == 3 ? d->nz : 1;
  P->nf = d->physics == MHX_PHYS_NAVIER_STOKES ? dim + 1 : 1;
  P->ne = 1; P->nnode = 1; P->nvert = 1;
  for (int a = 0; a < 3; ++a) {
    bool act = a < dim;
    P->nn[a] = act ? p * P->nel[a] + 1 : 1;
    P->nvx[a] = act ? P->nel[a] + 1 : 1;
    P->ne *= P->nel[a];
    P->nnode *= P->nn[a];
    P->nvert *= P->nvx[a];
  }
  P->ndof = P->nnode * P->nf;
  // reference AD size selection (assemblyManager_construct.hpp:128-171)
  int nd = P->nb * P->nf;
  P->ad_size = nd <= 2 ? 2 : nd <= 4 ? 4 : nd <= 8 ? 8 : nd <= 16 ? 16 : nd <= 18 ? 18 : nd <= 24 ? 24 : nd <= 32 ? 32 : 64;

  P->k0.assign(P->ne, 1.0);
  P->src.assign(P->ne, 0.0);
  if (d->k0) std::copy(d->k0, d->k0 + P->ne, P->k0.begin());
  if (d->src_elem) std::copy(d->src_elem, d->src_elem + P->ne, P->src.begin());
  set_functions(*P, *d);

  // vertices
  double lo[3] = {d->xmin, d->ymin, d->zmin}, hi[3] = {d->xmax, d->ymax, d->zmax};
  P->vx.assign(P->nvert * 3, 0.0);
  for (long v = 0; v < P->nvert; ++v) {
    long c[3] = {v % P->nvx[0], (v / P->nvx[0]) % P->nvx[1], v / ((long)P->nvx[0] * P->nvx[1])};
    bool interior = true;
    for (int a = 0; a < dim; ++a) if (c[a] == 0 || c[a] == P->nvx[a] - 1) interior = false;
    for (int a = 0; a < dim; ++a) {
      double h = (hi[a] - lo[a]) / P->nel[a];
      double x = lo[a] + c[a] * h;
      if (interior && d->perturb != 0.0)
        x += d->perturb * h * (2.0 * rng_u01(d->perturb_seed, (uint64_t)v * 3 + a) - 1.0);
      P->vx[v * 3 + a] = x;
    }
  }
  // connectivity, lexicographic local order; per-field element offsets
  // f*nb + j (the role of wkset->offsets(var, dof))
  const int nf = P->nf;
  P->lids.assign(P->ne * nf * P->nb, 0);
  for (long e = 0; e < P->ne; ++e) {
    long c[3];
    P->elem_coords(e, c);
    for (int j = 0; j < P->nb; ++j) {
      int jl[3] = {j % (p + 1), (j / (p + 1)) % (p + 1), dim == 3 ? j / ((p + 1) * (p + 1)) : 0};
      long g[3];
      for (int a = 0; a < 3; ++a) g[a] = a < dim ? p * c[a] + jl[a] : 0;
      const long node = g[0] + P->nn[0] * (g[1] + (long)P->nn[1] * g[2]);
      for (int f = 0; f < nf; ++f) P->lids[(e * nf + f) * P->nb + j] = (int)(node * nf + f);
    }
  }
  // 1-D rules and tensor tables
  gauss_legendre(P->nq1, P->qx1, P->qw1);
  std::vector<double> L1(P->nq1 * (p + 1)), D1(P->nq1 * (p + 1)), N1(P->nq1 * 2), M1(P->nq1 * 2);
  for (int q = 0; q < P->nq1; ++q) {
    lagrange1d(p, P->qx1[q], &L1[q * (p + 1)], &D1[q * (p + 1)]);
    lagrange1d(1, P->qx1[q], &N1[q * 2], &M1[q * 2]);
  }
  int nb = P->nb, nq = P->nq, nq1 = P->nq1;
  P->bref.assign(nb * nq, 0.0);
  P->gref.assign(nb * nq * dim, 0.0);
  P->qwref.assign(nq, 0.0);
  P->vbas.assign(P->nv * nq, 0.0);
  P->vgrad.assign(P->nv * nq * dim, 0.0);
  for (int q = 0; q < nq; ++q) {
    int qi[3] = {q % nq1, (q / nq1) % nq1, dim == 3 ? q / (nq1 * nq1) : 0};
    double w = 1.0;
    for (int a = 0; a < dim; ++a) w *= P->qw1[qi[a]];
    P->qwref[q] = w;
    for (int j = 0; j < nb; ++j) {
      int jl[3] = {j % (p + 1), (j / (p + 1)) % (p + 1), dim == 3 ? j / ((p + 1) * (p + 1)) : 0};
      double b = 1.0;
      for (int a = 0; a < dim; ++a) b *= L1[qi[a] * (p + 1) + jl[a]];
      P->bref[j * nq + q] = b;
      for (int c = 0; c < dim; ++c) {
        double g = 1.0;
        for (int a = 0; a < dim; ++a) g *= (a == c ? D1 : L1)[qi[a] * (p + 1) + jl[a]];
        P->gref[(j * nq + q) * dim + c] = g;
      }
    }
    for (int v = 0; v < P->nv; ++v) {
      int vl[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
      double b = 1.0;
      for (int a = 0; a < dim; ++a) b *= N1[qi[a] * 2 + vl[a]];
      P->vbas[v * nq + q] = b;
      for (int c = 0; c < dim; ++c) {
        double g = 1.0;
        for (int a = 0; a < dim; ++a) g *= (a == c ? M1 : N1)[qi[a] * 2 + vl[a]];
        P->vgrad[(v * nq + q) * dim + c] = g;
      }
    }
  }
  // CRS graph: union over elements of (gid_i x all gids of element), sorted
  // (linearAlgebraInterface_construct.hpp:243-265).
  std::vector<std::vector<int>> rows(P->ndof);
  const int nbd = nb * nf;
  for (long e = 0; e < P->ne; ++e) {
    const int* L = &P->lids[e * nbd];
    for (int i = 0; i < nbd; ++i)
      for (int j = 0; j < nbd; ++j) rows[L[i]].push_back(L[j]);
  }
  P->rowptr.assign(P->ndof + 1, 0);
  for (long r = 0; r < P->ndof; ++r) {
    auto& R = rows[r];
    std::sort(R.begin(), R.end());
    R.erase(std::unique(R.begin(), R.end()), R.end());
    P->rowptr[r + 1] = P->rowptr[r] + (int)R.size();
  }
  P->colidx.resize(P->rowptr[P->ndof]);
  for (long r = 0; r < P->ndof; ++r)
    std::copy(rows[r].begin(), rows[r].end(), P->colidx.begin() + P->rowptr[r]);
  // strong Dirichlet DOFs on the requested sides (setDirichletData,
  // discretizationInterface_dof.hpp:643-736) and their values.
  // Navier-Stokes: velocity component c on dbc_sides | ns_lid_sides |
  // ns_comp_sides[c] (lid value on the lid sides, 0 elsewhere), pressure
  // optionally pinned at node 0.
  P->fixed.assign(P->ndof, 0);
  P->dbcval.assign(P->ndof, 0.0);
  const bool ns = d->physics == MHX_PHYS_NAVIER_STOKES;
  const uint32_t vel_sides = d->dbc_sides | (ns ? d->ns_lid_sides : 0u);
  for (long node = 0; node < P->nnode; ++node) {
    long c[3] = {node % P->nn[0], (node / P->nn[0]) % P->nn[1], node / ((long)P->nn[0] * P->nn[1])};
    uint32_t on = 0;  // sides this node lies on
    for (int a = 0; a < dim; ++a) {
      if (c[a] == 0) on |= 1u << (2 * a);
      if (c[a] == P->nn[a] - 1) on |= 1u << (2 * a + 1);
    }
    for (int f = 0; f < nf; ++f) {
      const long g = node * nf + f;
      if (!ns) {
        if (on & d->dbc_sides) {
          P->fixed[g] = 1;
          P->dbcval[g] = d->dbc_values ? d->dbc_values[g] : d->dbc_const;
        }
      } else if (f < dim) {
        if (on & (vel_sides | d->ns_comp_sides[f])) {
          P->fixed[g] = 1;
          P->dbcval[g] = (on & d->ns_lid_sides) ? d->ns_lid_u[f] : 0.0;
        }
      } else if (d->ns_pin_pressure && node == 0) {
        P->fixed[g] = 1;
      }
    }
  }
  return P;
}

void Problem::vertex_of_elem(long e, int v, double* X) const {
  long c[3];
  elem_coords(e, c);
  long vc[3] = {c[0] + (v & 1), c[1] + ((v >> 1) & 1), dim == 3 ? c[2] + ((v >> 2) & 1) : 0};
  long vid = vc[0] + nvx[0] * (vc[1] + (long)nvx[1] * vc[2]);
  for (int a = 0; a < 3; ++a) X[a] = vx[vid * 3 + a];
}

// Physical integration data for one element: wts = w*|det J|, physical
// gradients J^{-T} grad_ref, and physical point coordinates
// (getPhysicalVolumetricBasis / getPhysicalIntegrationData,
//  discretizationInterface_basis.hpp:382-615, _integration.hpp:218-265).
struct ElemGeom {
  std::vector<double> wts, grad, x;  // [nq], [nb*nq*dim], [nq*3]
};

static void elem_geometry(const Problem& P, long e, ElemGeom& G) {
  int dim = P.dim, nq = P.nq, nb = P.nb, nv = P.nv;
  G.wts.assign(nq, 0.0);
  G.grad.assign(nb * nq * dim, 0.0);
  G.x.assign(nq * 3, 0.0);
  std::vector<double> X(nv * 3);
  for (int v = 0; v < nv; ++v) P.vertex_of_elem(e, v, &X[v * 3]);
  for (int q = 0; q < nq; ++q) {
    double J[3][3] = {{0}};
    for (int v = 0; v < nv; ++v) {
      for (int a = 0; a < dim; ++a) {
        G.x[q * 3 + a] += P.vbas[v * nq + q] * X[v * 3 + a];
        for (int b = 0; b < dim; ++b) J[a][b] += X[v * 3 + a] * P.vgrad[(v * nq + q) * dim + b];
      }
    }
    double Ji[3][3] = {{0}};
    double det;
    if (dim == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      Ji[0][0] = J[1][1] / det; Ji[0][1] = -J[0][1] / det;
      Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
    } else {
      det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
            J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
      Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
      Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
      Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
      Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
      Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
      Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
      Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
      Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    }
    G.wts[q] = P.qwref[q] * std::fabs(det);
    for (int j = 0; j < nb; ++j)
      for (int a = 0; a < dim; ++a) {
        double s = 0.0;
        for (int b = 0; b < dim; ++b) s += Ji[b][a] * P.gref[(j * nq + q) * dim + b];
        G.grad[(j * nq + q) * dim + a] = s;
      }
  }
}

// thermal source f(x) = src_e + sum_m amp_m prod_d s(w_md x_d)
static double source_at(const Problem& P, long e, const double* x) {
  double f = P.src[e];
  for (int m = 0; m < P.desc.n_modes; ++m) {
    double t = P.desc.mode_amp[m];
    for (int a = 0; a < P.dim; ++a) {
      double w = P.desc.mode_freq[m][a];
      if (w != 0.0) t = t * std::sin(w * x[a]);
    }
    f += t;
  }
  return f;
}

static inline long find_slot(const Problem& P, int row, int col) {
  const int* b = &P.colidx[P.rowptr[row]];
  const int* en = &P.colidx[P.rowptr[row + 1]];
  const int* it = std::lower_bound(b, en, col);
  return (long)(it - &P.colidx[0]);
}

// assembleJacRes / assembleRes for the thermal module.
// Structure follows assemblyManager_jacres.hpp:121-630 (sequential loop over
// worksets of `workset size` elements: gather -> seed -> physics -> scatter),
// with the field evaluation of workset.cpp:979-1115, the thermal volume
// residual of thermal.cpp:71-165 (steady, no advection) and the scatter of
// assemblyManager_scatter.hpp:164-278; then dofConstraints
// (assemblyManager_constraints.hpp:241-270).
// With P.threads > 1 the worksets run on that many OpenMP threads and the
// scatter adds atomically, as the reference's Kokkos-OpenMP build does
// (assemblyManager_scatter.hpp:227, 261: Kokkos::atomic_add); the sums'
// order then varies from run to run.  One thread (the default) keeps the
// sequential order.
template <class EvalT>
static void assemble_thermal(const Problem& P, const double* u, bool want_jac, double* res, double* jv) {
  const int ws = P.desc.workset_size > 0 ? P.desc.workset_size : 100;
  const int nb = P.nb, nq = P.nq, dim = P.dim;
  const long nws = (P.ne + ws - 1) / ws;
  const bool par = P.threads > 1;
#pragma omp parallel num_threads(P.threads) if (par)
  {
  std::vector<EvalT> uAD, T, dT;
  ElemGeom G;
#pragma omp for schedule(dynamic, 1)
  for (long wsi = 0; wsi < nws; ++wsi) {
    const long e0 = wsi * ws;
    long nel = std::min<long>(ws, P.ne - e0);
    uAD.assign(nel * nb, EvalT(0.0));
    // performGather + computeSolnSteadySeeded
    for (long le = 0; le < nel; ++le)
      for (int j = 0; j < nb; ++j) {
        double uv = u[P.lids[(e0 + le) * nb + j]];
        uAD[le * nb + j] = want_jac ? Seed<EvalT>::make(j, uv) : EvalT(uv);
      }
    for (long le = 0; le < nel; ++le) {
      long e = e0 + le;
      elem_geometry(P, e, G);
      // the evaluated solution of a Butcher stage: alpha_u u + beta_u per DOF
      const EvalT* uE = &uAD[le * nb];
      std::vector<EvalT> ue;
      if (!P.beta_u.empty()) {
        ue.resize(nb);
        for (int j = 0; j < nb; ++j) ue[j] = P.alpha_u * uAD[le * nb + j] + P.beta_u[P.lids[e * nb + j]];
        uE = ue.data();
      }
      // evaluateSolutionField: field = u_0 b_0, then += u_dof b_dof
      T.assign(nq, EvalT(0.0));
      dT.assign(nq * dim, EvalT(0.0));
      for (int q = 0; q < nq; ++q) {
        T[q] = uE[0] * P.bref[q];
        for (int a = 0; a < dim; ++a) dT[q * dim + a] = uE[0] * G.grad[(0 * nq + q) * dim + a];
        for (int j = 1; j < nb; ++j) {
          T[q] += uE[j] * P.bref[j * nq + q];
          for (int a = 0; a < dim; ++a) dT[q * dim + a] += uE[j] * G.grad[(j * nq + q) * dim + a];
        }
      }
      // time derivative field "T_t": the seeded nodal u_dot = alpha_t u + beta_t
      // evaluated like T (zero when steady)
      std::vector<EvalT> Tt(nq, EvalT(0.0));
      if (P.bdf > 0) {
        for (int j = 0; j < nb; ++j) {
          EvalT ud = uAD[le * nb + j] * P.alpha_t + P.beta_t[P.lids[e * nb + j]];
          for (int q = 0; q < nq; ++q) Tt[q] += ud * P.bref[j * nq + q];
        }
      }
      // FunctionManager::evaluate("thermal source" / "thermal diffusion" /
      // "density" / "specific heat" / "bx" / "by" / "bz")
      std::vector<EvalT> source(nq), diff(nq), rhocp(nq), bx, by, bz;
      bool adv = false;
      if (P.fm_active) {
        FElem<EvalT> el;
        el.nq = nq;
        el.T = T.data();
        el.dT = dT.data();
        el.Tt = P.bdf > 0 ? Tt.data() : nullptr;
        el.x = G.x.data();
        el.dim = dim;
        el.time = P.desc.fn_time;
        el.k0 = P.k0[e];
        el.src = P.src[e];
        auto fn = [&](const char* name, double dflt) {
          return P.fm.find(name) >= 0 ? P.fm.values<EvalT>(name, el) : std::vector<EvalT>(nq, EvalT(dflt));
        };
        diff = fn("thermal diffusion", 1.0);
        source = fn("thermal source", 0.0);
        const std::vector<EvalT> rho = fn("density", 1.0), cp = fn("specific heat", 1.0);
        for (int q = 0; q < nq; ++q) rhocp[q] = rho[q] * cp[q];
        adv = P.fm.find("advection x") >= 0 || P.fm.find("advection y") >= 0 || P.fm.find("advection z") >= 0;
        if (adv) {
          bx = fn("advection x", 0.0);
          by = fn("advection y", 0.0);
          bz = fn("advection z", 0.0);
        }
      } else {
        const double rho = P.desc.th_density > 0.0 ? P.desc.th_density : 1.0;
        const double cp = P.desc.th_specific_heat > 0.0 ? P.desc.th_specific_heat : 1.0;
        for (int q = 0; q < nq; ++q) {
          source[q] = EvalT(source_at(P, e, &G.x[q * 3]));
          EvalT TT = T[q] * T[q];
          EvalT onep = 1.0 + TT * P.desc.diff_u2;
          diff[q] = P.k0[e] * (P.desc.diff_u2 != 0.0 ? onep : EvalT(1.0));
          rhocp[q] = EvalT(rho * cp);
        }
      }
      // "Thermal volume resid 3D part 1": (rho cp dT/dt - source) w phi +
      // diff grad T . grad phi w (+ b . grad T w phi with advection)
      std::vector<EvalT> r(nb, EvalT(0.0));
      for (int i = 0; i < nb; ++i) {
        for (int q = 0; q < nq; ++q) {
          r[i] += (rhocp[q] * Tt[q] - source[q]) * G.wts[q] * P.bref[i * nq + q];
          for (int a = 0; a < dim; ++a)
            r[i] += diff[q] * dT[q * dim + a] * G.wts[q] * G.grad[(i * nq + q) * dim + a];
          if (adv) {
            EvalT bg = bx[q] * dT[q * dim];
            if (dim > 1) bg += by[q] * dT[q * dim + 1];
            if (dim > 2) bg += bz[q] * dT[q * dim + 2];
            r[i] += bg * G.wts[q] * P.bref[i * nq + q];
          }
        }
      }
      // scatter: res -= r.val(); J(row, cols) += r.dx(col); skip fixed rows
      const int* L = &P.lids[e * nb];
      for (int i = 0; i < nb; ++i) {
        int row = L[i];
        if (P.fixed[row]) continue;
        if (res) {
          const double v = -val(r[i]);
          if (par) {
#pragma omp atomic
            res[row] += v;
          } else {
            res[row] += v;
          }
        }
        if (want_jac && jv)
          for (int j = 0; j < nb; ++j) {
            const double v = dx(r[i], j);
            double* slot = &jv[find_slot(P, row, L[j])];
            if (par) {
#pragma omp atomic
              *slot += v;
            } else {
              *slot += v;
            }
          }
      }
    }
  }
  }  // omp parallel
  if (want_jac && jv)
    for (long g = 0; g < P.ndof; ++g)
      if (P.fixed[g]) jv[find_slot(P, (int)g, (int)g)] = 1.0;
}

// thermal<EvalT>::boundaryResidual for "Neumann" sides (thermal.cpp:210-222),
// run on the boundary groups after the volume groups (assemblyManager_jacres.hpp
// :389-401): res(e, dof) += -g w_side phi_dof at the side cubature points,
// scattered as res_glob -= res; g = q(x) . n does not depend on u, so the
// Jacobian is unchanged.  Side cubature: the tensor Gauss rule of the volume on
// the face axes; side weights w * |surface Jacobian| of the element's Q1 map
// restricted to the face; n the outward unit normal.
static void neumann_residual(const Problem& P, double* res) {
  const mhx_problem_desc& d = P.desc;
  if (!res || !d.neumann_sides || d.n_nq_modes <= 0) return;
  const int dim = P.dim, p = P.p, nq1 = P.nq1, nb = P.nb;
  std::vector<double> L(p + 1), dL(p + 1);
  for (int s = 0; s < 2 * dim; ++s) {
    if (!(d.neumann_sides & (1u << s))) continue;
    const int a = s / 2;
    const bool hi = s & 1;
    const int ta[2] = {a == 0 ? 1 : 0, a == 2 ? 1 : 2};  // tangential axes (the first only in 2-D)
    const int nt = dim - 1;
    for (long e = 0; e < P.ne; ++e) {
      long ec[3];
      P.elem_coords(e, ec);
      if (ec[a] != (hi ? P.nel[a] - 1 : 0)) continue;
      std::vector<double> X(P.nv * 3);
      for (int v = 0; v < P.nv; ++v) P.vertex_of_elem(e, v, &X[v * 3]);
      const int nqs = nt == 1 ? nq1 : nq1 * nq1;
      for (int qs = 0; qs < nqs; ++qs) {
        double xi[3] = {0, 0, 0}, w = 1.0;
        xi[a] = hi ? 1.0 : -1.0;
        const int qq[2] = {qs % nq1, qs / nq1};
        for (int t = 0; t < nt; ++t) {
          xi[ta[t]] = P.qx1[qq[t]];
          w *= P.qw1[qq[t]];
        }
        // Q1 map: x and the tangent vectors dx/dxi_t
        double x[3] = {0, 0, 0}, tg[2][3] = {{0, 0, 0}, {0, 0, 0}};
        for (int v = 0; v < P.nv; ++v) {
          double N = 1.0, dN[3] = {1.0, 1.0, 1.0};
          for (int b = 0; b < dim; ++b) {
            const double sgn = ((v >> b) & 1) ? 1.0 : -1.0, f = 0.5 * (1.0 + sgn * xi[b]);
            N *= f;
            for (int c = 0; c < dim; ++c) dN[c] *= (c == b) ? 0.5 * sgn : f;
          }
          for (int c = 0; c < 3; ++c) {
            x[c] += N * X[v * 3 + c];
            for (int t = 0; t < nt; ++t) tg[t][c] += dN[ta[t]] * X[v * 3 + c];
          }
        }
        double nrm[3] = {0, 0, 0}, area;
        if (dim == 2) {
          area = std::sqrt(tg[0][0] * tg[0][0] + tg[0][1] * tg[0][1]);
          nrm[0] = tg[0][1] / area;
          nrm[1] = -tg[0][0] / area;
        } else {
          nrm[0] = tg[0][1] * tg[1][2] - tg[0][2] * tg[1][1];
          nrm[1] = tg[0][2] * tg[1][0] - tg[0][0] * tg[1][2];
          nrm[2] = tg[0][0] * tg[1][1] - tg[0][1] * tg[1][0];
          area = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
          for (int c = 0; c < 3; ++c) nrm[c] /= area;
        }
        const double outward = hi ? 1.0 : -1.0;
        if (nrm[a] * outward < 0) for (int c = 0; c < 3; ++c) nrm[c] = -nrm[c];
        w *= area;
        // thermal: g = q . n (one field); Navier-Stokes: the traction q_c
        // of each velocity component (navierstokes.cpp:853-1000)
        const int ncomp = P.nf == 1 ? 1 : dim;
        double gc[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < d.n_nq_modes; ++m) {
          double prod = 1.0;
          for (int b = 0; b < dim; ++b) prod *= std::sin(d.nq_freq[m][b] * x[b] + d.nq_phase[m][b]);
          for (int c = 0; c < dim; ++c) {
            if (P.nf == 1) gc[0] += d.nq_amp[m][c] * prod * nrm[c];
            else gc[c] += d.nq_amp[m][c] * prod;
          }
        }
        // basis of every element DOF at the side point
        std::vector<double> Lb(dim * (p + 1));
        for (int b = 0; b < dim; ++b) {
          lagrange1d(p, xi[b], L.data(), dL.data());
          for (int k = 0; k <= p; ++k) Lb[b * (p + 1) + k] = L[k];
        }
        for (int j = 0; j < nb; ++j) {
          const int jl[3] = {j % (p + 1), (j / (p + 1)) % (p + 1), dim == 3 ? j / ((p + 1) * (p + 1)) : 0};
          double phi = 1.0;
          for (int b = 0; b < dim; ++b) phi *= Lb[b * (p + 1) + jl[b]];
          for (int c = 0; c < ncomp; ++c) {
            const int row = P.lids[(e * P.nf + c) * nb + j];
            if (P.fixed[row] || phi == 0.0) continue;
            res[row] += -(-gc[c] * w * phi);
          }
        }
      }
    }
  }
}

// navierstokes<EvalT>::computeTau (navierstokes.cpp:1053-1081), steady
// (C3 = 0): tau = 1 / sqrt((C1 nu / h^2)^2 + (C2 |u| / h)^2), where |u|^2 is
// replaced by its square root only when it exceeds 1e-12, as there.
template <class EvalT>
static EvalT ns_tau(double visc, const EvalT* vel, int dim, double h) {
  const double C1 = 4.0, C2 = 2.0;
  EvalT nvel = vel[0] * vel[0];
  for (int a = 1; a < dim; ++a) nvel = nvel + vel[a] * vel[a];
  if (val(nvel) > 1e-12) nvel = dsqrt(nvel);
  const double t1 = (C1 * visc / h / h) * (C1 * visc / h / h);
  EvalT t2 = (C2 * nvel / h) * (C2 * nvel / h);
  EvalT tau = t2 + t1;
  return 1.0 / dsqrt(tau);
}

// assembleJacRes / assembleRes for the navier stokes module: the workset loop
// as for thermal, the fields ux, uy, [uz], pr and their gradients
// (workset.cpp:979-1115), and navierstokes<EvalT>::volumeResidual for the
// steady case without energy coupling (navierstokes.cpp:252-494 in 2-D,
// :494-~850 in 3-D): Galerkin momentum and continuity, SUPG (useSUPG) and
// PSPG (usePSPG) with the element size of Workset::getElementSize
// (workset.cpp:2699-2712), h = (sum_q wts)^(1/dim).
template <class EvalT>
static void assemble_ns(const Problem& P, const double* u, bool want_jac, double* res, double* jv) {
  const int ws = P.desc.workset_size > 0 ? P.desc.workset_size : 100;
  const int nb = P.nb, nq = P.nq, dim = P.dim, nf = P.nf, nbd = nb * nf;
  const double rho = P.desc.ns_rho, visc = P.desc.ns_mu;
  const double* force = P.desc.ns_force;
  std::vector<EvalT> uAD, U, dU;
  ElemGeom G;
  for (long e0 = 0; e0 < P.ne; e0 += ws) {
    const long nel = std::min<long>(ws, P.ne - e0);
    uAD.assign(nel * nbd, EvalT(0.0));
    for (long le = 0; le < nel; ++le)
      for (int k = 0; k < nbd; ++k) {
        const double uv = u[P.lids[(e0 + le) * nbd + k]];
        uAD[le * nbd + k] = want_jac ? Seed<EvalT>::make(k, uv) : EvalT(uv);
      }
    for (long le = 0; le < nel; ++le) {
      const long e = e0 + le;
      elem_geometry(P, e, G);
      // fields U[f][q] and gradients dU[(f*nq + q)*dim + a]
      U.assign(nf * nq, EvalT(0.0));
      dU.assign(nf * nq * dim, EvalT(0.0));
      for (int f = 0; f < nf; ++f)
        for (int q = 0; q < nq; ++q) {
          const EvalT* uf = &uAD[le * nbd + f * nb];
          U[f * nq + q] = uf[0] * P.bref[q];
          for (int a = 0; a < dim; ++a) dU[(f * nq + q) * dim + a] = uf[0] * G.grad[(0 * nq + q) * dim + a];
          for (int j = 1; j < nb; ++j) {
            U[f * nq + q] += uf[j] * P.bref[j * nq + q];
            for (int a = 0; a < dim; ++a) dU[(f * nq + q) * dim + a] += uf[j] * G.grad[(j * nq + q) * dim + a];
          }
        }
      double vol = 0.0;
      for (int q = 0; q < nq; ++q) vol += G.wts[q];
      const double h = std::pow(vol, 1.0 / (double)dim);
      const int pf = dim;  // pressure field
      std::vector<EvalT> r(nbd, EvalT(0.0));
      auto gradU = [&](int f, int q, int a) -> const EvalT& { return dU[(f * nq + q) * dim + a]; };
      for (int q = 0; q < nq; ++q) {
        const double w = G.wts[q];
        EvalT vel[3] = {EvalT(0.0), EvalT(0.0), EvalT(0.0)};
        for (int a = 0; a < dim; ++a) vel[a] = U[a * nq + q];
        const EvalT& pr = U[pf * nq + q];
        const bool stab = P.desc.ns_supg || P.desc.ns_pspg;
        const EvalT tau = stab ? ns_tau<EvalT>(visc, vel, dim, h) : EvalT(0.0);
        // momentum residual R_a = rho (u . grad u_a) + d_a p - rho f_a (steady)
        EvalT conv[3], stabres[3];
        for (int a = 0; a < dim; ++a) {
          conv[a] = vel[0] * gradU(a, q, 0);
          for (int c = 1; c < dim; ++c) conv[a] = conv[a] + vel[c] * gradU(a, q, c);
          stabres[a] = rho * conv[a] + gradU(pf, q, a) - rho * force[a];
        }
        for (int a = 0; a < dim; ++a) {
          EvalT Fl[3];
          for (int c = 0; c < dim; ++c) {
            Fl[c] = visc * gradU(a, q, c);
            if (c == a) Fl[c] = Fl[c] - pr;
            Fl[c] = Fl[c] * w;
          }
          EvalT F = (conv[a] - force[a]) * (rho * w);
          for (int i = 0; i < nb; ++i) {
            EvalT t = F * P.bref[i * nq + q];
            for (int c = 0; c < dim; ++c) t += Fl[c] * G.grad[(i * nq + q) * dim + c];
            r[a * nb + i] += t;
          }
          if (P.desc.ns_supg) {
            for (int i = 0; i < nb; ++i) {
              EvalT t(0.0);
              for (int c = 0; c < dim; ++c) t += tau * stabres[a] * vel[c] * w * G.grad[(i * nq + q) * dim + c];
              r[a * nb + i] += t;
            }
          }
        }
        // continuity: div u, and PSPG tau/rho R . grad q
        EvalT divu = gradU(0, q, 0);
        for (int c = 1; c < dim; ++c) divu = divu + gradU(c, q, c);
        divu = divu * w;
        for (int i = 0; i < nb; ++i) r[pf * nb + i] += divu * P.bref[i * nq + q];
        if (P.desc.ns_pspg) {
          EvalT S[3];
          for (int c = 0; c < dim; ++c) S[c] = stabres[c] * (tau * w / rho);
          for (int i = 0; i < nb; ++i) {
            EvalT t(0.0);
            for (int c = 0; c < dim; ++c) t += S[c] * G.grad[(i * nq + q) * dim + c];
            r[pf * nb + i] += t;
          }
        }
      }
      // scatter (assemblyManager_scatter.hpp:164-278)
      const int* L = &P.lids[e * nbd];
      for (int i = 0; i < nbd; ++i) {
        const int row = L[i];
        if (P.fixed[row]) continue;
        if (res) res[row] += -val(r[i]);
        if (want_jac && jv)
          for (int j = 0; j < nbd; ++j) jv[find_slot(P, row, L[j])] += dx(r[i], j);
      }
    }
  }
  if (want_jac && jv)
    for (long g = 0; g < P.ndof; ++g)
      if (P.fixed[g]) jv[find_slot(P, (int)g, (int)g)] = 1.0;
}

// weak Dirichlet data g(x) = wd_const + wd_lin . x + sum_m amp prod sin(freq x + phase)
static double weak_dbc_data(const mhx_problem_desc& d, int dim, const double* x) {
  double g = d.wd_const;
  for (int b = 0; b < dim; ++b) g += d.wd_lin[b] * x[b];
  for (int m = 0; m < d.n_wd_modes; ++m) {
    double prod = d.wd_amp[m];
    for (int b = 0; b < dim; ++b) prod *= std::sin(d.wd_freq[m][b] * x[b] + d.wd_phase[m][b]);
    g += prod;
  }
  return g;
}

// thermal<EvalT>::boundaryResidual for "weak Dirichlet" sides (thermal.cpp:
// 226-281), on the boundary groups after the volume groups
// (assemblyManager_jacres.hpp:389-401):
//   res(e,dof) += epen/h k (T - g) w phi - k (grad T . n) w phi - sf k (T - g) w (grad phi . n)
// with epen = 10, h = (sum of the side weights)^(1/(dim-1))
// (Workset::getSideElementSize, workset.cpp:2719-2733), sf = "form_param"
// (thermal.cpp:35), k the thermal diffusion at T.  Side cubature and normals
// as in neumann_residual; grad phi through the element's Q1 map at the side
// point.  Scatter as the volume: res_glob -= res, J += d res / d u, strong
// Dirichlet rows skipped.
template <class EvalT>
static void nitsche_boundary(const Problem& P, const double* u, bool want_jac, double* res, double* jv) {
  const mhx_problem_desc& d = P.desc;
  if (!d.wdbc_sides) return;
  const int dim = P.dim, p = P.p, nq1 = P.nq1, nb = P.nb;
  const double epen = 10.0;
  const double sf = d.wdbc_form == 1 ? -1.0 : d.wdbc_form == 2 ? 0.0 : 1.0;
  std::vector<double> L(p + 1), dL(p + 1);
  for (int s = 0; s < 2 * dim; ++s) {
    if (!(d.wdbc_sides & (1u << s))) continue;
    const int a = s / 2;
    const bool hi = s & 1;
    const int ta[2] = {a == 0 ? 1 : 0, a == 2 ? 1 : 2};
    const int nt = dim - 1;
    const int nqs = nt == 1 ? nq1 : nq1 * nq1;
    for (long e = 0; e < P.ne; ++e) {
      long ec[3];
      P.elem_coords(e, ec);
      if (ec[a] != (hi ? P.nel[a] - 1 : 0)) continue;
      std::vector<double> X(P.nv * 3);
      for (int v = 0; v < P.nv; ++v) P.vertex_of_elem(e, v, &X[v * 3]);
      std::vector<EvalT> uAD(nb);
      for (int j = 0; j < nb; ++j) {
        const double uv = u[P.lids[e * nb + j]];
        uAD[j] = want_jac ? Seed<EvalT>::make(j, uv) : EvalT(uv);
      }
      // side points: weights, normals, physical coordinates, basis and gradients
      std::vector<double> wq(nqs), nq(nqs * 3), xq(nqs * 3), phi(nb * nqs), gphi(nb * nqs * dim);
      for (int qs = 0; qs < nqs; ++qs) {
        double xi[3] = {0, 0, 0}, w = 1.0;
        xi[a] = hi ? 1.0 : -1.0;
        const int qq[2] = {qs % nq1, qs / nq1};
        for (int t = 0; t < nt; ++t) {
          xi[ta[t]] = P.qx1[qq[t]];
          w *= P.qw1[qq[t]];
        }
        double x[3] = {0, 0, 0}, tg[2][3] = {{0, 0, 0}, {0, 0, 0}}, J[3][3] = {{0}};
        for (int v = 0; v < P.nv; ++v) {
          double N = 1.0, dN[3] = {1.0, 1.0, 1.0};
          for (int b = 0; b < dim; ++b) {
            const double sgn = ((v >> b) & 1) ? 1.0 : -1.0, f = 0.5 * (1.0 + sgn * xi[b]);
            N *= f;
            for (int c = 0; c < dim; ++c) dN[c] *= (c == b) ? 0.5 * sgn : f;
          }
          for (int c = 0; c < 3; ++c) {
            x[c] += N * X[v * 3 + c];
            for (int t = 0; t < nt; ++t) tg[t][c] += dN[ta[t]] * X[v * 3 + c];
          }
          for (int r = 0; r < dim; ++r)
            for (int b = 0; b < dim; ++b) J[r][b] += X[v * 3 + r] * dN[b];
        }
        double nrm[3] = {0, 0, 0}, area;
        if (dim == 2) {
          area = std::sqrt(tg[0][0] * tg[0][0] + tg[0][1] * tg[0][1]);
          nrm[0] = tg[0][1] / area;
          nrm[1] = -tg[0][0] / area;
        } else {
          nrm[0] = tg[0][1] * tg[1][2] - tg[0][2] * tg[1][1];
          nrm[1] = tg[0][2] * tg[1][0] - tg[0][0] * tg[1][2];
          nrm[2] = tg[0][0] * tg[1][1] - tg[0][1] * tg[1][0];
          area = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
          for (int c = 0; c < 3; ++c) nrm[c] /= area;
        }
        if (nrm[a] * (hi ? 1.0 : -1.0) < 0) for (int c = 0; c < 3; ++c) nrm[c] = -nrm[c];
        wq[qs] = w * area;
        for (int c = 0; c < 3; ++c) {
          nq[qs * 3 + c] = nrm[c];
          xq[qs * 3 + c] = x[c];
        }
        double Ji[3][3] = {{0}};
        if (dim == 2) {
          const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
          Ji[0][0] = J[1][1] / det; Ji[0][1] = -J[0][1] / det;
          Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
        } else {
          const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                             J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
          Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
          Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
          Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
          Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
          Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
          Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
          Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
          Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
          Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        }
        std::vector<double> Lb(dim * (p + 1)), Db(dim * (p + 1));
        for (int b = 0; b < dim; ++b) {
          lagrange1d(p, xi[b], L.data(), dL.data());
          for (int k = 0; k <= p; ++k) {
            Lb[b * (p + 1) + k] = L[k];
            Db[b * (p + 1) + k] = dL[k];
          }
        }
        for (int j = 0; j < nb; ++j) {
          const int jl[3] = {j % (p + 1), (j / (p + 1)) % (p + 1), dim == 3 ? j / ((p + 1) * (p + 1)) : 0};
          double ph = 1.0, gr[3];
          for (int b = 0; b < dim; ++b) ph *= Lb[b * (p + 1) + jl[b]];
          for (int c = 0; c < dim; ++c) {
            gr[c] = 1.0;
            for (int b = 0; b < dim; ++b) gr[c] *= (b == c ? Db : Lb)[b * (p + 1) + jl[b]];
          }
          phi[j * nqs + qs] = ph;
          for (int c = 0; c < dim; ++c) {
            double sg = 0.0;
            for (int b = 0; b < dim; ++b) sg += Ji[b][c] * gr[b];
            gphi[(j * nqs + qs) * dim + c] = sg;
          }
        }
      }
      double vol = 0.0;
      for (int qs = 0; qs < nqs; ++qs) vol += wq[qs];
      const double h = std::pow(vol, 1.0 / ((double)dim - 1.0));
      std::vector<EvalT> r(nb, EvalT(0.0));
      for (int qs = 0; qs < nqs; ++qs) {
        // evaluateSolutionField on the side: field = u_0 b_0, then += u_dof b_dof
        EvalT T = uAD[0] * phi[qs];
        EvalT dT[3];
        for (int c = 0; c < dim; ++c) dT[c] = uAD[0] * gphi[qs * dim + c];
        for (int j = 1; j < nb; ++j) {
          T += uAD[j] * phi[j * nqs + qs];
          for (int c = 0; c < dim; ++c) dT[c] += uAD[j] * gphi[(j * nqs + qs) * dim + c];
        }
        EvalT TT = T * T;
        EvalT onep = 1.0 + TT * d.diff_u2;
        EvalT k = P.k0[e] * (d.diff_u2 != 0.0 ? onep : EvalT(1.0));
        const double g = weak_dbc_data(d, dim, &xq[qs * 3]);
        EvalT dTn = dT[0] * nq[qs * 3];
        for (int c = 1; c < dim; ++c) dTn += dT[c] * nq[qs * 3 + c];
        const double w = wq[qs];
        for (int i = 0; i < nb; ++i) {
          double gpn = 0.0;
          for (int c = 0; c < dim; ++c) gpn += gphi[(i * nqs + qs) * dim + c] * nq[qs * 3 + c];
          r[i] += epen / h * k * (T - g) * w * phi[i * nqs + qs];
          r[i] += -k * dTn * w * phi[i * nqs + qs];
          r[i] += -sf * k * (T - g) * w * gpn;
        }
      }
      const int* Lr = &P.lids[e * nb];
      for (int i = 0; i < nb; ++i) {
        const int row = Lr[i];
        if (P.fixed[row]) continue;
        if (res) res[row] += -val(r[i]);
        if (want_jac && jv)
          for (int j = 0; j < nb; ++j) jv[find_slot(P, row, Lr[j])] += dx(r[i], j);
      }
    }
  }
}

static void assemble(const Problem& P, const double* u, bool want_jac, double* res, double* jv) {
  const bool ns = P.desc.physics == MHX_PHYS_NAVIER_STOKES;
  auto run = [&](auto tag) {
    using EvalT = typename decltype(tag)::type;
    if (ns) {
      assemble_ns<EvalT>(P, u, want_jac, res, jv);
    } else {
      assemble_thermal<EvalT>(P, u, want_jac, res, jv);
      nitsche_boundary<EvalT>(P, u, want_jac, res, jv);  // boundary groups, after the volume groups
    }
  };
  struct D { using type = double; };
  if (!want_jac) {
    run(D{});
  } else {
    switch (P.ad_size) {
      case 2:  { struct T { using type = SFad<2>; }; run(T{}); break; }
      case 4:  { struct T { using type = SFad<4>; }; run(T{}); break; }
      case 8:  { struct T { using type = SFad<8>; }; run(T{}); break; }
      case 16: { struct T { using type = SFad<16>; }; run(T{}); break; }
      case 18: { struct T { using type = SFad<18>; }; run(T{}); break; }
      case 24: { struct T { using type = SFad<24>; }; run(T{}); break; }
      case 32: { struct T { using type = SFad<32>; }; run(T{}); break; }
      default: { struct T { using type = SFad<64>; }; run(T{}); break; }
    }
  }
  neumann_residual(P, res);  // boundary groups, after the volume groups
}

// ---------------------------------------------------------------------------
// Krylov
// ---------------------------------------------------------------------------
static void matvec(const Problem& P, const double* A, const double* x, double* y) {
#pragma omp parallel for num_threads(P.threads) if (P.threads > 1) schedule(static)
  for (long r = 0; r < P.ndof; ++r) {
    double s = 0.0;
    for (int k = P.rowptr[r]; k < P.rowptr[r + 1]; ++k) s += A[k] * x[P.colidx[k]];
    y[r] = s;
  }
}
// (with more than one thread the sum's order follows the thread split)
static int g_dot_threads = 1;
static double dot(long n, const double* a, const double* b) {
  double s = 0;
#pragma omp parallel for num_threads(g_dot_threads) if (g_dot_threads > 1) reduction(+ : s) schedule(static)
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// SolverManager::PCG (solverManager_solvers.hpp:812-905): Jacobi PCG,
// stopping on ||r||/||r0|| <= tol with r0 = b - A x0.
static int pcg(const Problem& P, const double* A, const double* b, double* x, double tol, int maxit, double* relres) {
  long n = P.ndof;
  std::vector<double> M(n), p(n, 0.0), q(n, 0.0), r(n), z(n, 0.0);
  for (long i = 0; i < n; ++i) M[i] = A[find_slot(P, (int)i, (int)i)];
  matvec(P, A, x, q.data());
  for (long i = 0; i < n; ++i) r[i] = b[i] - q[i];
  double rn = std::sqrt(dot(n, r.data(), r.data()));
  double r0 = rn;
  double rho = 1.0, rho1;
  int it = 0;
  if (r0 == 0.0) { if (relres) *relres = 0.0; return 0; }
  while (it < maxit && rn / r0 > tol) {
    for (long i = 0; i < n; ++i) z[i] = r[i] / M[i];
    rho1 = rho;
    rho = dot(n, r.data(), z.data());
    if (it == 0) p = z;
    else { double beta = rho / rho1; for (long i = 0; i < n; ++i) p[i] = z[i] + beta * p[i]; }
    matvec(P, A, p.data(), q.data());
    double alpha = rho / dot(n, p.data(), q.data());
    for (long i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    rn = std::sqrt(dot(n, r.data(), r.data()));
    ++it;
  }
  if (relres) *relres = rn / r0;
  return it;
}

// Restarted GMRES(m), right Jacobi preconditioning, modified Gram-Schmidt,
// Givens rotations; stops on ||b - A x||/||b|| <= tol (true residual checked
// at each restart).  Stands in for Belos Block GMRES + Ifpack2 Jacobi
// (linearAlgebraInterface_solvers.hpp:51-80,100-141), whose iterates are
// not reproducible outside Trilinos: only the solution is compared.
static int gmres(const Problem& P, const double* A, const double* b, double* x, int m, double tol, int maxit,
                 double* relres) {
  long n = P.ndof;
  std::vector<double> Minv(n), r(n), w(n), tmp(n);
  for (long i = 0; i < n; ++i) Minv[i] = 1.0 / A[find_slot(P, (int)i, (int)i)];
  double bn = std::sqrt(dot(n, b, b));
  if (bn == 0.0) { std::fill(x, x + n, 0.0); if (relres) *relres = 0.0; return 0; }
  std::vector<std::vector<double>> V(m + 1, std::vector<double>(n));
  std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1);
  int it = 0;
  double rel = 1.0;
  while (it < maxit) {
    matvec(P, A, x, tmp.data());
    for (long i = 0; i < n; ++i) r[i] = b[i] - tmp[i];
    double beta = std::sqrt(dot(n, r.data(), r.data()));
    rel = beta / bn;
    if (rel <= tol) break;
    for (long i = 0; i < n; ++i) V[0][i] = r[i] / beta;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (; k < m && it < maxit; ++k, ++it) {
      const int nt = P.threads;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
      for (long i = 0; i < n; ++i) tmp[i] = Minv[i] * V[k][i];
      matvec(P, A, tmp.data(), w.data());
      for (int j = 0; j <= k; ++j) {
        double h = dot(n, w.data(), V[j].data());
        H[j * m + k] = h;
        const double* vj = V[j].data();
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
        for (long i = 0; i < n; ++i) w[i] -= h * vj[i];
      }
      double hn = std::sqrt(dot(n, w.data(), w.data()));
      H[(k + 1) * m + k] = hn;
      double* vk1 = V[k + 1].data();
      if (hn != 0.0) {
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
        for (long i = 0; i < n; ++i) vk1[i] = w[i] / hn;
      }
      for (int j = 0; j < k; ++j) {
        double a = H[j * m + k], bb = H[(j + 1) * m + k];
        H[j * m + k] = cs[j] * a + sn[j] * bb;
        H[(j + 1) * m + k] = -sn[j] * a + cs[j] * bb;
      }
      double a = H[k * m + k], bb = H[(k + 1) * m + k];
      double den = std::sqrt(a * a + bb * bb);
      cs[k] = a / den; sn[k] = bb / den;
      H[k * m + k] = den; H[(k + 1) * m + k] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      if (std::fabs(g[k + 1]) / bn <= tol * 0.5) { ++k; ++it; break; }
    }
    std::vector<double> y(k);
    for (int j = k - 1; j >= 0; --j) {
      double s = g[j];
      for (int l = j + 1; l < k; ++l) s -= H[j * m + l] * y[l];
      y[j] = s / H[j * m + j];
    }
    std::fill(tmp.begin(), tmp.end(), 0.0);
    for (int j = 0; j < k; ++j) for (long i = 0; i < n; ++i) tmp[i] += y[j] * V[j][i];
    for (long i = 0; i < n; ++i) x[i] += Minv[i] * tmp[i];
  }
  matvec(P, A, x, tmp.data());
  for (long i = 0; i < n; ++i) r[i] = b[i] - tmp[i];
  rel = std::sqrt(dot(n, r.data(), r.data())) / bn;
  if (relres) *relres = rel;
  return it;
}

static int linear_solve(const Problem& P, const double* A, const double* b, double* x, const mhx_krylov_opts* o,
                        mhx_solve_stats* st) {
  g_dot_threads = P.threads;
  double rel = 0.0;
  int it;
  if (o->solver == MHX_KRYLOV_CG) it = pcg(P, A, b, x, o->tol, o->max_iters, &rel);
  else it = gmres(P, A, b, x, o->restart > 0 ? o->restart : 50, o->tol, o->max_iters, &rel);
  if (st) { st->iters += it; st->rel_res = rel; }
  return it;
}

// SolverManager::nonlinearSolver (solverManager_solvers.hpp:290-643):
// residual-only assembly, inf-norm test relative to the first norm,
// optional backtracking, Jacobian assembly, linear solve, u += du.
static int newton(const Problem& P, double* u, const mhx_newton_opts* o, mhx_solve_stats* st) {
  long n = P.ndof;
  std::vector<double> res(n), du(n, 0.0), jv(P.colidx.size());
  double first = 0.0, alpha = 1.0;
  int it = 0;
  bool proceed = true;
  while (proceed) {
    std::fill(res.begin(), res.end(), 0.0);
    assemble(P, u, false, res.data(), nullptr);
    double nrm = 0.0;
    for (long i = 0; i < n; ++i) nrm = std::max(nrm, std::fabs(res[i]));
    double scaled;
    if (it == 0) { first = nrm; scaled = 1.0; } else scaled = nrm / first;
    if (st) { if (it == 0) st->res_norm0 = nrm; st->res_norm = nrm; }
    bool solve = true;
    if (o->backtracking && scaled > 1.1) {
      solve = false;
      alpha *= 0.5;
      for (long i = 0; i < n; ++i) u[i] -= alpha * du[i];
    } else if (!o->relative_off) {  // "use relative TOL" (solverManager_solvers.hpp:431-439)
      if (scaled < o->nl_tol || nrm < 1e-100) {
        solve = false;
        proceed = false;
      }
    } else if (o->use_abs_tol && nrm < (o->abs_tol > 0 ? o->abs_tol : std::min(o->nl_tol, 1e-6))) {
      solve = false;  // "use absolute TOL" (:441-443)
      proceed = false;
    }
    if (solve) {
      std::fill(jv.begin(), jv.end(), 0.0);
      std::vector<double> r2(n, 0.0);
      assemble(P, u, true, r2.data(), jv.data());
      std::fill(du.begin(), du.end(), 0.0);
      linear_solve(P, jv.data(), res.data(), du.data(), &o->linear, st);
      alpha = 1.0;
      for (long i = 0; i < n; ++i) u[i] += alpha * du[i];
    }
    ++it;
    if (it >= o->max_iters) proceed = false;
  }
  if (st) st->nl_iters = it;
  return 0;
}

// L2 / H1-seminorm errors at the volume cubature points
// (postprocessManager_error_estimation.hpp:173-330; sqrt at print,
//  postprocessManager_util.hpp:778-794).
static void l2_error(const Problem& P, const double* u, int nm, const double* amp, const double* freq, double* l2,
                     double* l2g) {
  double e0 = 0.0, e1 = 0.0;
  ElemGeom G;
  int nb = P.nb, nq = P.nq, dim = P.dim;
  for (long e = 0; e < P.ne; ++e) {
    elem_geometry(P, e, G);
    const int* L = &P.lids[e * nb];
    for (int q = 0; q < nq; ++q) {
      double Th = 0.0, gh[3] = {0, 0, 0};
      for (int j = 0; j < nb; ++j) {
        Th += u[L[j]] * P.bref[j * nq + q];
        for (int a = 0; a < dim; ++a) gh[a] += u[L[j]] * G.grad[(j * nq + q) * dim + a];
      }
      double Tt = 0.0, gt[3] = {0, 0, 0};
      const double* x = &G.x[q * 3];
      for (int m = 0; m < nm; ++m) {
        double s[3], c[3];
        for (int a = 0; a < dim; ++a) {
          double w = freq[m * 3 + a];
          s[a] = w != 0.0 ? std::sin(w * x[a]) : 1.0;
          c[a] = w != 0.0 ? w * std::cos(w * x[a]) : 0.0;
        }
        double t = amp[m];
        for (int a = 0; a < dim; ++a) t *= s[a];
        Tt += t;
        for (int a = 0; a < dim; ++a) {
          double g = amp[m];
          for (int b = 0; b < dim; ++b) g *= (a == b ? c[b] : s[b]);
          gt[a] += g;
        }
      }
      e0 += (Th - Tt) * (Th - Tt) * G.wts[q];
      for (int a = 0; a < dim; ++a) e1 += (gh[a] - gt[a]) * (gh[a] - gt[a]) * G.wts[q];
    }
  }
  *l2 = std::sqrt(e0);
  *l2g = std::sqrt(e1);
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C entry points used by the tests through ctypes.
// ---------------------------------------------------------------------------
extern "C" {

const char* orc_last_error(void) { return orc::g_err.c_str(); }

int orc_create(const mhx_problem_desc* d, void** out) {
  if (!d || !out || d->dim < 2 || d->dim > 3 || d->order < 1 || d->quadrature < 0 || d->nx < 1 || d->ny < 1 ||
      (d->dim == 3 && d->nz < 1)) {
    orc::g_err = "orc_create: bad description";
    return MHX_ERR_ARG;
  }
  if (d->physics != MHX_PHYS_THERMAL && d->physics != MHX_PHYS_NAVIER_STOKES) {
    orc::g_err = "orc_create: unknown physics";
    return MHX_ERR_UNSUPPORTED;
  }
  if (d->physics == MHX_PHYS_NAVIER_STOKES && !(d->ns_rho > 0.0 && d->ns_mu > 0.0)) {
    orc::g_err = "orc_create: density and viscosity must be positive";
    return MHX_ERR_ARG;
  }
  try {
    *out = orc::build_problem(d);
  } catch (const std::exception& ex) {
    orc::g_err = ex.what();
    return MHX_ERR_ARG;
  }
  return MHX_OK;
}

// FunctionManager::evaluate of one function at n points: each point its own
// element (one_element = 0), or all n points one element (1, for the element
// reductions).  pts[11 i ..]: T, T_t, grad(T)[x], [y], [z], x, y, z, t, k0,
// src; out[6 i ..]: value and the derivatives by T, grad(T)[x|y|z], T_t.
int orc_expr_eval(const mhx_problem_desc* d, const char* name, int64_t n, const double* pts, double* out,
                  int32_t one_element) {
  using E = orc::SFad<5>;
  try {
    orc::FManager fm;
    fm.have_k0 = fm.have_src = true;
    for (int i = 0; i < d->n_functions; ++i) fm.add(d->function_names[i], d->function_exprs[i]);
    fm.decompose();
    const int64_t per = one_element ? n : 1;
    for (int64_t s0 = 0; s0 < n; s0 += per) {
      std::vector<E> T(per), dT(per * 3), Tt(per);
      std::vector<double> x(per * 3);
      for (int64_t q = 0; q < per; ++q) {
        const double* p = pts + 11 * (s0 + q);
        T[q] = E(5, 0, p[0]);
        Tt[q] = E(5, 4, p[1]);
        for (int c = 0; c < 3; ++c) {
          dT[q * 3 + c] = E(5, 1 + c, p[2 + c]);
          x[q * 3 + c] = p[5 + c];
        }
      }
      const double* p0 = pts + 11 * s0;
      orc::FElem<E> el;
      el.nq = (int)per;
      el.T = T.data();
      el.dT = dT.data();
      el.Tt = Tt.data();
      el.x = x.data();
      el.dim = 3;
      el.time = p0[8];
      el.k0 = p0[9];
      el.src = p0[10];
      const std::vector<E> v = fm.values<E>(name, el);
      for (int64_t q = 0; q < per; ++q) {
        out[6 * (s0 + q)] = v[q].v;
        for (int k = 0; k < 5; ++k) out[6 * (s0 + q) + 1 + k] = v[q].d[k];
      }
    }
  } catch (const std::exception& ex) {
    orc::g_err = ex.what();
    return MHX_ERR_ARG;
  }
  return MHX_OK;
}
void orc_destroy(void* h) { delete (orc::Problem*)h; }

int orc_sizes(void* h, int64_t* ndof, int64_t* nnz, int64_t* ne, int32_t* nb, int32_t* nq, int32_t* ad) {
  auto* P = (orc::Problem*)h;
  *ndof = P->ndof; *nnz = (int64_t)P->colidx.size(); *ne = P->ne; *nb = P->nb; *nq = P->nq; *ad = P->ad_size;
  return MHX_OK;
}
int orc_pattern(void* h, int32_t* rowptr, int32_t* colidx) {
  auto* P = (orc::Problem*)h;
  std::copy(P->rowptr.begin(), P->rowptr.end(), rowptr);
  std::copy(P->colidx.begin(), P->colidx.end(), colidx);
  return MHX_OK;
}
int orc_dirichlet(void* h, int32_t* mask, double* vals) {
  auto* P = (orc::Problem*)h;
  for (long g = 0; g < P->ndof; ++g) { if (mask) mask[g] = P->fixed[g]; if (vals) vals[g] = P->dbcval[g]; }
  return MHX_OK;
}
int orc_set_dirichlet(void* h, double* u) {
  auto* P = (orc::Problem*)h;
  for (long g = 0; g < P->ndof; ++g) if (P->fixed[g]) u[g] = P->dbcval[g];
  return MHX_OK;
}
int orc_lids(void* h, int32_t* lids) {
  auto* P = (orc::Problem*)h;
  std::copy(P->lids.begin(), P->lids.end(), lids);
  return MHX_OK;
}
// res/jv are accumulated into (the caller zeroes them, as
// solverManager_solvers.hpp:347-354 does).
// threads of the thermal assembly (the OpenMP baseline); returns the count used
int orc_set_threads(void* h, int32_t n) {
  auto* P = (orc::Problem*)h;
  P->threads = n < 1 ? 1 : n;
  return P->threads;
}
int orc_assemble(void* h, const double* u, int32_t want_jac, double* res, double* jv) {
  orc::assemble(*(orc::Problem*)h, u, want_jac != 0, res, jv);
  return MHX_OK;
}
int orc_linear_solve(void* h, const double* A, const double* b, double* x, const mhx_krylov_opts* o,
                     mhx_solve_stats* st) {
  orc::linear_solve(*(orc::Problem*)h, A, b, x, o, st);
  return MHX_OK;
}
int orc_newton(void* h, double* u, const mhx_newton_opts* o, mhx_solve_stats* st) {
  return orc::newton(*(orc::Problem*)h, u, o, st);
}
// Time-step / Butcher-stage state (SolverManager::setBackwardDifference,
// solverManager_setup.hpp:443-500, weights without 1/dt; the stage terms of
// Workset::computeSolnTransientSeeded, workset.cpp:601-680):
//   timewt = 1/dt/b_s, alpha_t = BDF_0 timewt,
//   beta_t = (sum_k BDF_k u_prev[k-1]) timewt,
//   alpha_u = A_ss/b_s, beta_u = (1 - alpha_u) u_prev[0]
//                               + sum_{j<s} A_sj/b_j (u_stages[j] - u_prev[0]).
int orc_set_stage_state(void* h, double dt, int32_t order, const double* u_prev, int32_t nst, const double* A,
                        const double* b, int32_t stage, const double* u_stages) {
  auto* P = (orc::Problem*)h;
  static const double W[7][7] = {
      {0}, {1.0, -1.0}, {1.5, -2.0, 0.5}, {11.0 / 6.0, -3.0, 1.5, -1.0 / 3.0},
      {25.0 / 12.0, -4.0, 3.0, -4.0 / 3.0, 1.0 / 4.0},
      {137.0 / 60.0, -5.0, 5.0, -10.0 / 3.0, 75.0 / 60.0, -1.0 / 5.0},
      {147.0 / 60.0, -6.0, 15.0 / 2.0, -20.0 / 3.0, 225.0 / 60.0, -72.0 / 60.0, 1.0 / 6.0}};
  if (order < 0 || order > 6 || (order > 0 && !(dt > 0.0))) return MHX_ERR_ARG;
  P->bdf = order;
  P->beta_u.clear();
  P->alpha_u = 1.0;
  if (order == 0) {
    P->beta_t.clear();
    P->alpha_t = 0.0;
    return MHX_OK;
  }
  if (nst < 1 || stage < 0 || stage >= nst || b[stage] == 0.0) return MHX_ERR_ARG;
  const double timewt = 1.0 / dt / b[stage];
  P->alpha_t = W[order][0] * timewt;
  P->beta_t.assign(P->ndof, 0.0);
  for (long g = 0; g < P->ndof; ++g) {
    double bt = 0.0;
    for (int s = 1; s <= order; ++s) bt += W[order][s] * u_prev[(long)(s - 1) * P->ndof + g];
    P->beta_t[g] = bt * timewt;
  }
  const double ass = A[stage * nst + stage], alpha_u = ass / b[stage];
  if (nst > 1 || ass != 1.0 || b[stage] != 1.0) {
    P->alpha_u = alpha_u;
    P->beta_u.assign(P->ndof, 0.0);
    for (long g = 0; g < P->ndof; ++g) {
      const double u0 = u_prev[g];
      double bu = (1.0 - alpha_u) * u0;
      for (int j = 0; j < stage; ++j) bu += A[stage * nst + j] / b[j] * (u_stages[(long)j * P->ndof + g] - u0);
      P->beta_u[g] = bu;
    }
  }
  return MHX_OK;
}
int orc_set_time_state(void* h, double dt, int32_t order, const double* u_prev) {
  const double one = 1.0;
  return orc_set_stage_state(h, dt, order, u_prev, 1, &one, &one, 0, nullptr);
}
// coefficient update between steps: source modes, density, specific heat and
// the constant Dirichlet value (the mesh and sides stay)
int orc_update(void* h, const mhx_problem_desc* d) {
  auto* P = (orc::Problem*)h;
  try {
    if (d->k0) std::copy(d->k0, d->k0 + P->ne, P->k0.begin());
    if (d->src_elem) std::copy(d->src_elem, d->src_elem + P->ne, P->src.begin());
    orc::set_functions(*P, *d);
  } catch (const std::exception& ex) {
    orc::g_err = ex.what();
    return MHX_ERR_ARG;
  }
  P->desc.fn_time = d->fn_time;
  P->desc.n_modes = d->n_modes;
  std::copy(d->mode_amp, d->mode_amp + MHX_MAX_MODES, P->desc.mode_amp);
  for (int m = 0; m < MHX_MAX_MODES; ++m)
    for (int a = 0; a < 3; ++a) P->desc.mode_freq[m][a] = d->mode_freq[m][a];
  P->desc.th_density = d->th_density;
  P->desc.th_specific_heat = d->th_specific_heat;
  P->desc.wd_const = d->wd_const;
  P->desc.n_wd_modes = d->n_wd_modes;
  for (int b = 0; b < 3; ++b) P->desc.wd_lin[b] = d->wd_lin[b];
  for (int m = 0; m < MHX_MAX_NQ_MODES; ++m) {
    P->desc.wd_amp[m] = d->wd_amp[m];
    for (int b = 0; b < 3; ++b) {
      P->desc.wd_freq[m][b] = d->wd_freq[m][b];
      P->desc.wd_phase[m][b] = d->wd_phase[m][b];
    }
  }
  if (!d->dbc_values) {
    P->desc.dbc_const = d->dbc_const;
    for (long g = 0; g < P->ndof; ++g)
      if (P->fixed[g]) P->dbcval[g] = d->dbc_const;
  }
  return MHX_OK;
}
int orc_l2_error(void* h, const double* u, int32_t nm, const double* amp, const double* freq, double* l2,
                 double* l2g) {
  orc::l2_error(*(orc::Problem*)h, u, nm, amp, freq, l2, l2g);
  return MHX_OK;
}
// Physical basis values/gradients of element e at its cubature points, in
// the oracle's lexicographic local order: phi[nb*nq], grad[nb*nq*dim], wts[nq], x[nq*3].
int orc_elem_basis(void* h, int64_t e, double* phi, double* grad, double* wts, double* x) {
  auto* P = (orc::Problem*)h;
  orc::ElemGeom G;
  orc::elem_geometry(*P, e, G);
  if (phi) std::copy(P->bref.begin(), P->bref.end(), phi);
  if (grad) std::copy(G.grad.begin(), G.grad.end(), grad);
  if (wts) std::copy(G.wts.begin(), G.wts.end(), wts);
  if (x) std::copy(G.x.begin(), G.x.end(), x);
  return MHX_OK;
}
int orc_gauss(int32_t n, double* x, double* w) {
  std::vector<double> X, W;
  orc::gauss_legendre(n, X, W);
  std::copy(X.begin(), X.end(), x);
  std::copy(W.begin(), W.end(), w);
  return MHX_OK;
}
int orc_dof_coords(void* h, double* xyz) {
  // node (ix,iy,iz) of the lattice mapped through the Q1 map of an element containing it
  auto* P = (orc::Problem*)h;
  int p = P->p;
  for (long g = 0; g < P->ndof; ++g) {
    const long node = g / P->nf;
    long c[3] = {node % P->nn[0], (node / P->nn[0]) % P->nn[1], node / ((long)P->nn[0] * P->nn[1])};
    long ec[3] = {0, 0, 0};
    double xi[3] = {0, 0, 0};
    for (int a = 0; a < P->dim; ++a) {
      ec[a] = std::min<long>(c[a] / p, P->nel[a] - 1);
      xi[a] = -1.0 + 2.0 * (double)(c[a] - p * ec[a]) / p;
    }
    long e = ec[0] + P->nel[0] * (ec[1] + (long)P->nel[1] * ec[2]);
    double X[3] = {0, 0, 0};
    for (int v = 0; v < P->nv; ++v) {
      double V[3];
      P->vertex_of_elem(e, v, V);
      double b = 1.0;
      for (int a = 0; a < P->dim; ++a) b *= ((v >> a) & 1) ? 0.5 * (1 + xi[a]) : 0.5 * (1 - xi[a]);
      for (int a = 0; a < 3; ++a) X[a] += b * V[a];
    }
    for (int a = 0; a < 3; ++a) xyz[g * 3 + a] = X[a];
  }
  return MHX_OK;
}

}  // extern "C"
