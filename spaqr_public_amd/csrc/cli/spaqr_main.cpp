// spaQR command-line driver on the MI355X library: the flags, defaults and
// flow of the reference driver (spaQR.cpp:25-323) over the C ABI.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/spaqr.h"

namespace {

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Csc {
  int m = 0, n = 0;
  std::vector<int> p, i;
  std::vector<double> x;
};

// Matrix Market coordinate reader (real/integer/pattern, general/symmetric),
// duplicates summed, sorted CSC output (mmio.hpp:174-230 semantics).
Csc read_mm_sparse(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::string line;
  std::getline(f, line);
  if (line.rfind("%%MatrixMarket", 0) != 0) throw std::runtime_error("not a Matrix Market file: " + path);
  // symmetric / hermitian mirror the off-diagonal entries, skew-symmetric
  // mirrors them negated (mmio.hpp:200-212)
  const bool skew = line.find("skew-symmetric") != std::string::npos;
  const bool symmetric = !skew && (line.find("symmetric") != std::string::npos ||
                                   line.find("hermitian") != std::string::npos);
  const bool pattern = line.find("pattern") != std::string::npos;
  if (line.find("coordinate") == std::string::npos) throw std::runtime_error("expected a coordinate matrix: " + path);
  while (std::getline(f, line))
    if (!line.empty() && line[0] != '%') break;
  long m, n, nnz;
  std::istringstream(line) >> m >> n >> nnz;
  std::vector<std::vector<std::pair<int, double>>> cols(n);
  for (long k = 0; k < nnz; ++k) {
    long r, c;
    double v = 1.0;
    f >> r >> c;
    if (!pattern) f >> v;
    cols[c - 1].push_back({(int)(r - 1), v});
    if (symmetric && r != c) cols[r - 1].push_back({(int)(c - 1), v});
    if (skew && r != c) cols[r - 1].push_back({(int)(c - 1), -v});
  }
  Csc A;
  A.m = (int)m;
  A.n = (int)n;
  A.p.push_back(0);
  for (auto& col : cols) {
    std::sort(col.begin(), col.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t q = 0; q < col.size(); ++q) {
      if (!A.i.empty() && (int)A.i.size() > A.p.back() && A.i.back() == col[q].first) { A.x.back() += col[q].second; continue; }
      A.i.push_back(col[q].first);
      A.x.push_back(col[q].second);
    }
    A.p.push_back((int)A.i.size());
  }
  return A;
}

// Matrix Market dense array reader (mmio.hpp:235-320): a general real
// array, entries column by column, comment and blank lines skipped; the
// number of entries must be rows * cols.
std::vector<double> read_mm_dense(const std::string& path, int* rows, int* cols) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::string line;
  std::getline(f, line);
  if (line.rfind("%%MatrixMarket", 0) != 0) throw std::runtime_error("not a Matrix Market file: " + path);
  if (line.find("array") == std::string::npos) throw std::runtime_error("expected a dense array: " + path);
  if (line.find("general") == std::string::npos) throw std::runtime_error("expected a general array: " + path);
  if (line.find("pattern") != std::string::npos || line.find("complex") != std::string::npos)
    throw std::runtime_error("expected a real array: " + path);
  while (std::getline(f, line))
    if (!line.empty() && line[0] != '%') break;
  int m = 0, n = 1;
  std::istringstream(line) >> m >> n;
  std::vector<double> v;
  v.reserve((size_t)m * n);
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '%') continue;
    std::istringstream ls(line);
    double x;
    if (ls >> x) v.push_back(x);
  }
  if (v.size() != (size_t)m * n)
    throw std::runtime_error(path + ": " + std::to_string(v.size()) + " entries for a " + std::to_string(m) + " x " +
                             std::to_string(n) + " array");
  *rows = m;
  *cols = n;
  return v;
}

Csc transpose(const Csc& A) {
  Csc T;
  T.m = A.n;
  T.n = A.m;
  T.p.assign(A.m + 1, 0);
  for (int r : A.i) T.p[r + 1]++;
  for (int r = 0; r < A.m; ++r) T.p[r + 1] += T.p[r];
  T.i.resize(A.i.size());
  T.x.resize(A.x.size());
  std::vector<int> pos(T.p.begin(), T.p.end() - 1);
  for (int c = 0; c < A.n; ++c)
    for (int q = A.p[c]; q < A.p[c + 1]; ++q) {
      const int d = pos[A.i[q]]++;
      T.i[d] = c;
      T.x[d] = A.x[q];
    }
  return T;
}

void spmv(const Csc& A, const std::vector<double>& x, std::vector<double>& y) {
  y.assign(A.m, 0.0);
  for (int c = 0; c < A.n; ++c)
    for (int q = A.p[c]; q < A.p[c + 1]; ++q) y[A.i[q]] += A.x[q] * x[c];
}
// y = A^T x
void spmv_t(const Csc& A, const std::vector<double>& x, std::vector<double>& y) {
  y.assign(A.n, 0.0);
  for (int c = 0; c < A.n; ++c)
    for (int q = A.p[c]; q < A.p[c + 1]; ++q) y[c] += A.x[q] * x[A.i[q]];
}
double nrm(const std::vector<double>& v) {
  double s = 0;
  for (double a : v) s += a * a;
  return std::sqrt(s);
}

void check(int rc, const char* what) {
  if (rc != SPAQR_OK) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, spaqr_last_error());
    std::exit(2);
  }
}

struct Opt {
  std::string name, def;
  bool flag;
};

void usage() {
  printf("spaQR - Sparsified QR for general sparse matrices (MI355X)\n"
         "Usage: spaqr -m <matrix.mm> [options]\n"
         "  --help                     Print help\n"
         "  -m, --matrix arg           Matrix file in matrix market format\n"
         "  -l, --lvl arg              Number of levels (default ceil(log2(ncols/64)))\n"
         "      --coordinates arg      Coordinates MM array file (d x N)\n"
         "  -n, --coordinates_n arg    Tensor grid n^d coordinates (default: -1)\n"
         "  -d, --coordinates_d arg    Tensor grid dimension (default: -1)\n"
         "      --parts arg            Row partition MM array file (m x 1) for a matrix without\n"
         "                             coordinates (default: the library's own partitioner)\n"
         "      --hsl arg              Use HSL bipartite matching (default: 0; unavailable)\n"
         "  -t, --tol arg              Tolerance (default: 1e-1)\n"
         "      --skip arg             Skip sparsification (default: 0)\n"
         "      --order arg            Order of the scheme (1, 1.5) (default: 1)\n"
         "      --scale arg            Do scaling (default: 1)\n"
         "  -i, --iterations arg       Iterative solver iterations (default: 300)\n"
         "      --rhs arg              RHS in matrix market format\n"
         "      --res arg              Desired relative residual (default: 1e-12)\n"
         "      --useEigenLSCG arg     Eigen baseline (not available)\n"
         "      --useEigenQR arg       Eigen baseline (not available)\n"
         "      --useCholesky arg      Eigen baseline (not available)\n"
         "      --precision arg        fp64 or fp32 factor storage (default: fp64)\n"
         "      --device arg           HIP device (default: 0)\n"
         "      --profile arg          Write the factorization profile (JSON) to this file\n");
}

}  // namespace

static int run(int argc, char** argv) {
  // option table: long name -> (short, default)
  std::map<std::string, std::string> val = {
      {"coordinates_n", "-1"}, {"coordinates_d", "-1"}, {"hsl", "0"},        {"tol", "1e-1"},
      {"skip", "0"},           {"order", "1"},          {"scale", "1"},      {"iterations", "300"},
      {"res", "1e-12"},        {"useEigenLSCG", "0"},   {"useEigenQR", "0"}, {"useCholesky", "0"},
      {"precision", "fp64"},   {"device", "0"}};
  std::map<std::string, std::string> shortmap = {{"m", "matrix"}, {"l", "lvl"},   {"n", "coordinates_n"},
                                                 {"d", "coordinates_d"}, {"t", "tol"}, {"i", "iterations"}};
  const std::vector<std::string> known = {"matrix", "lvl", "coordinates", "coordinates_n", "coordinates_d", "hsl",
                                          "tol", "skip", "order", "scale", "iterations", "rhs", "res",
                                          "useEigenLSCG", "useEigenQR", "useCholesky", "precision", "device",
                                          "profile", "parts"};
  std::map<std::string, int> given;
  for (int a = 1; a < argc; ++a) {
    std::string s = argv[a], key, v;
    if (s == "--help") { usage(); return 0; }
    if (s.rfind("--", 0) == 0) {
      key = s.substr(2);
      auto eq = key.find('=');
      if (eq != std::string::npos) { v = key.substr(eq + 1); key = key.substr(0, eq); }
    } else if (s.size() >= 2 && s[0] == '-') {
      auto it = shortmap.find(s.substr(1, 1));
      if (it == shortmap.end()) { fprintf(stderr, "Option '%s' does not exist\n", s.c_str()); return 1; }
      key = it->second;
      if (s.size() > 2) v = s.substr(2);
    } else {
      fprintf(stderr, "Unexpected argument '%s'\n", s.c_str());
      return 1;
    }
    if (std::find(known.begin(), known.end(), key) == known.end()) {
      fprintf(stderr, "Option '%s' does not exist\n", key.c_str());
      return 1;
    }
    if (v.empty()) {
      if (a + 1 >= argc) { fprintf(stderr, "Option '%s' is missing an argument\n", key.c_str()); return 1; }
      v = argv[++a];
    }
    val[key] = v;
    given[key] = 1;
  }
  if (!given.count("matrix")) { printf("--matrix is mandatory\n"); return 0; }
  const std::string matrix = val["matrix"];
  const int cn = std::stoi(val["coordinates_n"]), cd = std::stoi(val["coordinates_d"]);
  if ((cn == -1 && cd >= 0) || (cd == -1 && cn >= 0)) {
    printf("cn and cd should be both provided, or none should be provided\n");
    return 1;
  }
  const bool geo_tensor = cn >= 0 && cd >= 0, geo_file = given.count("coordinates") > 0;
  if (std::stoi(val["useEigenLSCG"]) || std::stoi(val["useEigenQR"]) || std::stoi(val["useCholesky"])) {
    fprintf(stderr, "The Eigen baseline solvers are not part of this build.\n");
    return 2;
  }
  Csc A = read_mm_sparse(matrix);
  if (A.m < A.n) {
    printf(" <<< Warning!!! nrows < ncols. Finding QR on A.transpose() instead. \n");
    A = transpose(A);
  }
  const int nrows = A.m, ncols = A.n;
  printf("Matrix %s with %d rows,  %d columns loaded\n", matrix.c_str(), nrows, ncols);
  const int iterations = std::stoi(val["iterations"]);
  const double residual = std::stod(val["res"]);
  int nlevels;
  if (!given.count("lvl")) {
    printf("--Levels not provided\n");
    nlevels = (int)std::ceil(std::log2(ncols / 64));
    printf(" Levels set to ceil(log2(ncols/64)) =  %d\n", nlevels);
  } else {
    nlevels = std::stoi(val["lvl"]);
  }
  int scale = std::stoi(val["scale"]);
  if (nrows != ncols && scale == 0) {
    printf("Scaling necessary for rectangular matrices\nSetting scale to 1\n");
    scale = 1;
  }
  const double tol = std::stod(val["tol"]);
  const float order = std::stof(val["order"]);
  const int precision = val["precision"] == "fp32" ? SPAQR_FP32 : SPAQR_FP64;
  const double p0 = now();
  check(spaqr_colscale(A.m, A.n, A.p.data(), A.i.data(), A.x.data(), nullptr), "colscale");
  printf("Pre-process time: %g\n", now() - p0);
  const int skip = tol == 0 ? nlevels - 1 : std::stoi(val["skip"]);
  spaqr_tree* t = spaqr_tree_create(nlevels, skip, precision, std::stoi(val["device"]));
  if (!t) { fprintf(stderr, "tree creation failed: %s\n", spaqr_last_error()); return 2; }
  if (geo_tensor) {
    if ((long)std::pow(cn, cd) != ncols) {
      printf("Error: cn and cd where both provided, but cn^cd != N where A is NxN\n");
      return 1;
    }
    check(spaqr_set_tensor_coordinates(t, cn, cd), "coordinates");
    printf("Tensor coordinate matrix of size %d^%d built\n", cn, cd);
  } else if (geo_file) {
    int r, c;
    std::vector<double> X = read_mm_dense(val["coordinates"], &r, &c);
    printf("Coordinate file %dx%d loaded from %s\n", r, c, val["coordinates"].c_str());
    if (c != ncols) {
      printf("Error: coordinate file should hold a matrix of size d x N\n");
      return 1;
    }
    // stored column-major d x N == dof-major coordinates
    check(spaqr_set_coordinates(t, X.data(), r, c), "coordinates");
  } else if (given.count("parts")) {
    // no coordinates: the row partition of a file (m x 1 array of parts)
    int r, c;
    std::vector<double> P = read_mm_dense(val["parts"], &r, &c);
    if (r * c != A.m) {
      printf("Error: the parts file should hold one part per row (%d)\n", A.m);
      return 1;
    }
    std::vector<int> parts(P.begin(), P.end());
    check(spaqr_set_row_parts(t, parts.data(), A.m), "row parts");
    printf("Row partition loaded from %s\n", val["parts"].c_str());
  }
  check(spaqr_set_scale(t, scale), "set_scale");
  check(spaqr_set_tol(t, tol), "set_tol");
  check(spaqr_set_order(t, order), "set_order");
  check(spaqr_set_hsl(t, std::stoi(val["hsl"])), "set_hsl");
  if (nrows == ncols) check(spaqr_set_square(t, 1), "set_square");
  check(spaqr_set_verbose(t, 1), "set_verbose");
  check(spaqr_partition(t, A.m, A.n, A.p.data(), A.i.data(), A.x.data()), "partition");
  check(spaqr_assemble(t, A.m, A.n, A.p.data(), A.i.data(), A.x.data()), "assemble");
  check(spaqr_factorize(t), "factorize");
  {
    std::vector<double> b(nrows), x(ncols, 0.0), Ax;
    check(spaqr_random(nrows, 2021, b.data()), "random");
    const double t0 = now();
    if (nrows == ncols) {
      check(spaqr_solve(t, b.data(), x.data()), "solve");
      printf("<<<<tsolv=%.3e\n", now() - t0);
      printf("One-time solve (Random b):\n");
      spmv(A, x, Ax);
      for (int i = 0; i < nrows; ++i) Ax[i] -= b[i];
      printf("<<<<|(Ax-b)|/|b| : %.3e\n", nrm(Ax) / nrm(b));
    } else {
      // spaQR.cpp:277-283: one normal-equation apply of A^T b
      std::vector<double> Atb, r;
      spmv_t(A, b, Atb);
      check(spaqr_solve_nrml(t, Atb.data(), x.data()), "solve_nrml");
      printf("<<<<tsolv=%.3e\n", now() - t0);
      printf("One-time solve (Random b):\n");
      spmv(A, x, Ax);
      for (int i = 0; i < nrows; ++i) Ax[i] -= b[i];
      spmv_t(A, Ax, r);
      printf("<<<<|A'(Ax-b)|/|A'b| : %.3e\n", nrm(r) / nrm(Atb));
    }
  }
  {
    std::vector<double> b(nrows), x(ncols, 0.0), Ax;
    if (!given.count("rhs")) {
      check(spaqr_random(nrows, 2021, b.data()), "random");
    } else {
      int r, c;
      b = read_mm_dense(val["rhs"], &r, &c);  // vector_mmread (mmio.hpp:277-320)
      if (c != 1 || r != nrows) {
        fprintf(stderr, "--rhs: expected a %d x 1 array, got %d x %d\n", nrows, r, c);
        return 1;
      }
    }
    spaqr_matrix* M = spaqr_matrix_create(t, A.m, A.n, A.p.data(), A.i.data(), A.x.data());
    if (!M) { fprintf(stderr, "matrix upload failed: %s\n", spaqr_last_error()); return 2; }
    int it = 0;
    double est = 0;
    const double g0 = now();
    if (nrows == ncols) {
      check(spaqr_gmres(t, M, b.data(), x.data(), iterations, iterations, residual, 0, &it, &est), "gmres");
      const double g1 = now();
      spmv(A, x, Ax);
      for (int i = 0; i < nrows; ++i) Ax[i] -= b[i];
      // est is the solver's stopping quantity (is.cpp:217): the run
      // converged only if it fell below the requested residual
      if (est < residual) printf("GMRES converged!\n");
      else printf("GMRES failed to converge in %d iterations\n", it);
      printf("GMRES: #iterations: %d, residual |Ax-b|/|b|: %.3e\n", it, nrm(Ax) / nrm(b));
      printf("  GMRES: %.3e s.\n", g1 - g0);
      printf("<<<<GMRES=%d\n", it);
    } else {
      // spaQR.cpp:308-316
      check(spaqr_cgls(t, M, b.data(), x.data(), iterations, residual, 0, &it, &est), "cgls");
      const double g1 = now();
      std::vector<double> r, Atb;
      spmv(A, x, Ax);
      for (int i = 0; i < nrows; ++i) Ax[i] -= b[i];
      spmv_t(A, Ax, r);
      spmv_t(A, b, Atb);
      printf("CGLS: #iterations: %d, residual |A'(Ax-b)|/|A'(b)|: %.3e\n", it, nrm(r) / nrm(Atb));
      printf("  CGLS: %.3e s.\n", g1 - g0);
      printf("<<<<CGLS=%d\n", it);
    }
    spaqr_matrix_destroy(M);
  }
  if (given.count("profile")) {
    const int len = spaqr_profile_json(t, nullptr, 0);
    if (len < 0) check(len, "profile_json");
    std::string js((size_t)len + 1, '\0');
    check(spaqr_profile_json(t, &js[0], len + 1) < 0 ? SPAQR_E_ARG : SPAQR_OK, "profile_json");
    js.resize((size_t)len);
    std::ofstream pf(val["profile"]);
    if (!pf) throw std::runtime_error("cannot write " + val["profile"]);
    pf << js;
    printf("Profile written to %s\n", val["profile"].c_str());
  }
  spaqr_tree_destroy(t);
  return 0;
}

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {  // unreadable input files, bad flag values
    fprintf(stderr, "spaqr: %s\n", e.what());
    return 1;
  }
}
