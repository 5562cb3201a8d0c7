// mmio.hip — MatrixMarket files for the context's Jacobian and vectors.
//
// Replaces LinearAlgebraInterface::writeToFile
// (linearAlgebraInterface_util.hpp:316-331, "dump jacobian / residual /
// solution"), which writes through Tpetra::MatrixMarket::Writer:
// writeSparseFile for the CRS matrix ("coordinate real general", 1-based
// global indices, one entry per line) and writeDenseFile for vectors ("array
// real general", one column).  Readers place a file's entries into the
// context's CRS pattern, so a Jacobian dumped by a Trilinos run (after the DOF
// permutation of the binding) can be compared slot by slot.  Host code only.

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mhx_internal.hpp"

namespace mhx {

namespace {

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

void fetch(Ctx& c, const double* src, bool dev, size_t n, std::vector<double>& out) {
  out.resize(n);
  if (dev) {
    MHX_HIP(hipMemcpyAsync(out.data(), src, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    MHX_HIP(hipStreamSynchronize(c.stream));
  } else {
    std::copy(src, src + n, out.begin());
  }
}

// the CRS column indices live on the device only (built by k_fill_colidx)
std::vector<int32_t> colidx_host(Ctx& c) {
  std::vector<int32_t> ci((size_t)c.nnz);
  MHX_HIP(hipMemcpyAsync(ci.data(), c.colidx.p, ci.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  MHX_HIP(hipStreamSynchronize(c.stream));
  return ci;
}

void store(Ctx& c, const std::vector<double>& v, double* dst, bool dev) {
  if (dev) {
    MHX_HIP(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    MHX_HIP(hipStreamSynchronize(c.stream));
  } else {
    std::copy(v.begin(), v.end(), dst);
  }
}

// banner: "%%MatrixMarket matrix <format> <field> <symmetry>", lower-cased
struct Banner {
  std::string format, field, symmetry;
};

Banner read_banner(FILE* f, const char* path) {
  char line[1024];
  MHX_REQUIRE(std::fgets(line, sizeof line, f), MHX_ERR_ARG, std::string("empty MatrixMarket file ") + path);
  char tag[64] = {0}, obj[64] = {0}, fmt[64] = {0}, fld[64] = {0}, sym[64] = {0};
  const int k = std::sscanf(line, "%63s %63s %63s %63s %63s", tag, obj, fmt, fld, sym);
  auto lower = [](char* s) {
    for (; *s; ++s) *s = (char)std::tolower((unsigned char)*s);
  };
  lower(obj);
  lower(fmt);
  lower(fld);
  lower(sym);
  MHX_REQUIRE(k == 5 && std::strcmp(tag, "%%MatrixMarket") == 0 && std::strcmp(obj, "matrix") == 0, MHX_ERR_ARG,
              std::string("not a MatrixMarket matrix file: ") + path);
  MHX_REQUIRE(std::strcmp(fld, "real") == 0 || std::strcmp(fld, "integer") == 0, MHX_ERR_ARG,
              "MatrixMarket: only real or integer fields are read");
  MHX_REQUIRE(std::strcmp(sym, "general") == 0 || std::strcmp(sym, "symmetric") == 0, MHX_ERR_ARG,
              "MatrixMarket: only general or symmetric matrices are read");
  return Banner{fmt, fld, sym};
}

// first non-comment line
bool data_line(FILE* f, char* line, size_t n) {
  while (std::fgets(line, (int)n, f)) {
    const char* p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '%' || *p == '\n' || *p == '\r' || *p == 0) continue;
    return true;
  }
  return false;
}

}  // namespace

void write_matrix_market(Ctx& c, const char* path, const double* vals, bool dev) {
  std::vector<double> v;
  fetch(c, vals, dev, (size_t)c.nnz, v);
  const std::vector<int32_t> ci = colidx_host(c);
  File fl(path, "w");
  MHX_REQUIRE(fl.f, MHX_ERR_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
  const long n = c.mesh.ndof;
  std::fprintf(fl.f, "%%%%MatrixMarket matrix coordinate real general\n");
  std::fprintf(fl.f, "%% Jacobian of the mhx context, 1-based DOFs in lattice order\n");
  std::fprintf(fl.f, "%ld %ld %ld\n", n, n, (long)c.nnz);
  for (long r = 0; r < n; ++r)
    for (int k = c.h_rowptr[r]; k < c.h_rowptr[r + 1]; ++k)
      std::fprintf(fl.f, "%ld %d %.17g\n", r + 1, ci[k] + 1, v[k]);
  MHX_REQUIRE(std::fflush(fl.f) == 0 && !std::ferror(fl.f), MHX_ERR_ARG, std::string("write failed: ") + path);
}

void write_vector_market(Ctx& c, const char* path, const double* vec, bool dev) {
  std::vector<double> v;
  fetch(c, vec, dev, (size_t)c.mesh.ndof, v);
  File fl(path, "w");
  MHX_REQUIRE(fl.f, MHX_ERR_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
  std::fprintf(fl.f, "%%%%MatrixMarket matrix array real general\n");
  std::fprintf(fl.f, "%ld 1\n", (long)v.size());
  for (double x : v) std::fprintf(fl.f, "%.17g\n", x);
  MHX_REQUIRE(std::fflush(fl.f) == 0 && !std::ferror(fl.f), MHX_ERR_ARG, std::string("write failed: ") + path);
}

void read_matrix_market(Ctx& c, const char* path, double* vals, bool dev) {
  File fl(path, "r");
  MHX_REQUIRE(fl.f, MHX_ERR_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
  const Banner b = read_banner(fl.f, path);
  MHX_REQUIRE(b.format == "coordinate", MHX_ERR_ARG, "MatrixMarket: a sparse matrix must be in coordinate format");
  char line[1024];
  MHX_REQUIRE(data_line(fl.f, line, sizeof line), MHX_ERR_ARG, "MatrixMarket: missing size line");
  long nr = 0, nc = 0, ne = 0;
  MHX_REQUIRE(std::sscanf(line, "%ld %ld %ld", &nr, &nc, &ne) == 3, MHX_ERR_ARG, "MatrixMarket: bad size line");
  MHX_REQUIRE(nr == c.mesh.ndof && nc == c.mesh.ndof, MHX_ERR_ARG,
              "MatrixMarket: matrix size differs from the context's DOF count");
  const bool sym = b.symmetry == "symmetric";
  std::vector<double> v((size_t)c.nnz, 0.0);
  const std::vector<int32_t> ci = colidx_host(c);
  auto place = [&](long r, long col, double x) {
    const int32_t* base = ci.data();
    const int32_t* lo = base + c.h_rowptr[r];
    const int32_t* hi = base + c.h_rowptr[r + 1];
    const int32_t* it = std::lower_bound(lo, hi, (int32_t)col);
    MHX_REQUIRE(it != hi && *it == col, MHX_ERR_ARG,
                "MatrixMarket: entry (" + std::to_string(r + 1) + ", " + std::to_string(col + 1) +
                    ") lies outside the context's CRS pattern");
    v[it - base] += x;  // duplicates are summed
  };
  for (long k = 0; k < ne; ++k) {
    MHX_REQUIRE(data_line(fl.f, line, sizeof line), MHX_ERR_ARG, "MatrixMarket: fewer entries than announced");
    long r = 0, col = 0;
    double x = 0.0;
    MHX_REQUIRE(std::sscanf(line, "%ld %ld %lf", &r, &col, &x) == 3, MHX_ERR_ARG, "MatrixMarket: bad entry line");
    MHX_REQUIRE(r >= 1 && r <= nr && col >= 1 && col <= nc, MHX_ERR_ARG, "MatrixMarket: index out of range");
    place(r - 1, col - 1, x);
    if (sym && r != col) place(col - 1, r - 1, x);
  }
  store(c, v, vals, dev);
}

void read_vector_market(Ctx& c, const char* path, double* vec, bool dev) {
  File fl(path, "r");
  MHX_REQUIRE(fl.f, MHX_ERR_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
  const Banner b = read_banner(fl.f, path);
  char line[1024];
  MHX_REQUIRE(data_line(fl.f, line, sizeof line), MHX_ERR_ARG, "MatrixMarket: missing size line");
  const long n = c.mesh.ndof;
  std::vector<double> v((size_t)n, 0.0);
  if (b.format == "array") {
    long nr = 0, nc = 0;
    MHX_REQUIRE(std::sscanf(line, "%ld %ld", &nr, &nc) == 2 && nr == n && nc == 1, MHX_ERR_ARG,
                "MatrixMarket: expected an n x 1 array with n the context's DOF count");
    for (long i = 0; i < n; ++i) {
      MHX_REQUIRE(data_line(fl.f, line, sizeof line) && std::sscanf(line, "%lf", &v[i]) == 1, MHX_ERR_ARG,
                  "MatrixMarket: fewer values than announced");
    }
  } else {
    long nr = 0, nc = 0, ne = 0;
    MHX_REQUIRE(std::sscanf(line, "%ld %ld %ld", &nr, &nc, &ne) == 3 && nr == n && nc == 1, MHX_ERR_ARG,
                "MatrixMarket: expected an n x 1 matrix with n the context's DOF count");
    for (long k = 0; k < ne; ++k) {
      long r = 0, col = 0;
      double x = 0.0;
      MHX_REQUIRE(data_line(fl.f, line, sizeof line) && std::sscanf(line, "%ld %ld %lf", &r, &col, &x) == 3,
                  MHX_ERR_ARG, "MatrixMarket: bad entry line");
      MHX_REQUIRE(r >= 1 && r <= n && col == 1, MHX_ERR_ARG, "MatrixMarket: index out of range");
      v[r - 1] += x;
    }
  }
  store(c, v, vec, dev);
}

}  // namespace mhx
