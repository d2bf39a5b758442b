// Closed-form element operators of any order, in the layout of
// operator_data/p{p}.dat (element_init, mod_mesh.f90:253-263), and a writer
// for that format.
//
//  * x1D: Legendre-Gauss-Lobatto nodes, the end points -1, 1 and the roots of
//    P_p' (Newton on P_p' with P_p'' from Legendre's equation, started at the
//    Chebyshev-Gauss-Lobatto points, in extended precision).
//  * D1D(i,j) = l_j'(x_i): P_p(x_i) / (P_p(x_j) (x_i - x_j)) off the diagonal,
//    -p(p+1)/4 and p(p+1)/4 at the two end points, 0 elsewhere on the
//    diagonal (the LGL differentiation matrix; Kopriva, Implementing Spectral
//    Methods for PDEs, 2009, ch. 3).
//  * Lift_mat(:,:,:,f) = I (x) l-/+ (x) I along the face normal, with
//    l-/+(x) = sum_{n=0..p} (2n+1)/2 P_n(x) P_n(-/+1), the exact-mass-matrix
//    lift of a face value (faces -y, +x, +y, -x, -z, +z).
// For p <= 15 these reproduce the stored files to their rounding (the tests
// hold them to it); orders above 15, which have no files, use them directly.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "dgk_internal.hpp"

namespace dgk {

namespace {

// P_n(x) and P_{n-1}(x) by the three-term recurrence.
void legendre(int n, long double x, long double &pn, long double &pnm1) {
  long double p0 = 1.0L, p1 = x;
  if (n == 0) {
    pn = 1.0L;
    pnm1 = 0.0L;
    return;
  }
  for (int k = 1; k < n; ++k) {
    const long double p2 = ((2.0L * k + 1.0L) * x * p1 - k * p0) / (k + 1.0L);
    p0 = p1;
    p1 = p2;
  }
  pn = p1;
  pnm1 = p0;
}

}  // namespace

int make_operators(int p, std::vector<double> &x1D, std::vector<double> &D,
                   std::vector<double> &Lift) {
  if (p < 1) return set_error(DGK_ERR_ARG, "polynomial order must be >= 1");
  const int N = p, Nq = p + 1, Np = Nq * Nq * Nq;
  const long double pi = 3.141592653589793238462643383279502884L;
  std::vector<long double> x(Nq);
  x[0] = -1.0L;
  x[N] = 1.0L;
  for (int j = 1; j < N; ++j) {
    long double t = -std::cos(pi * j / N);
    for (int it = 0; it < 100; ++it) {
      long double pn, pnm1;
      legendre(N, t, pn, pnm1);
      const long double d1 = N * (t * pn - pnm1) / (t * t - 1.0L);       // P_N'
      const long double d2 = (2.0L * t * d1 - N * (N + 1.0L) * pn) / (1.0L - t * t);
      const long double dt = d1 / d2;
      t -= dt;
      if (std::fabs(dt) < 1e-19L) break;
    }
    x[j] = t;
  }
  for (int j = 1; j < N - j; ++j) {  // exact antisymmetry
    const long double s = 0.5L * (x[N - j] - x[j]);
    x[j] = -s;
    x[N - j] = s;
  }
  if (N % 2 == 0) x[N / 2] = 0.0L;

  std::vector<long double> PN(Nq);
  for (int i = 0; i < Nq; ++i) {
    long double pnm1;
    legendre(N, x[i], PN[i], pnm1);
  }
  x1D.assign(Nq, 0.0);
  for (int i = 0; i < Nq; ++i) x1D[i] = (double)x[i];
  D.assign((size_t)Nq * Nq, 0.0);
  for (int j = 0; j < Nq; ++j)
    for (int i = 0; i < Nq; ++i) {
      long double v;
      if (i != j) v = PN[i] / (PN[j] * (x[i] - x[j]));
      else v = i == 0 ? -N * (N + 1.0L) / 4.0L : i == N ? N * (N + 1.0L) / 4.0L : 0.0L;
      D[i + (size_t)j * Nq] = (double)v;
    }
  // l-(x_a), l+(x_a)
  std::vector<long double> lm(Nq, 0.0L), lp(Nq, 0.0L);
  for (int a = 0; a < Nq; ++a) {
    long double p0 = 1.0L, p1 = x[a];
    for (int n = 0; n <= N; ++n) {
      const long double Pn = n == 0 ? 1.0L : n == 1 ? x[a] : p1;
      lm[a] += (2.0L * n + 1.0L) / 2.0L * Pn * (n % 2 ? -1.0L : 1.0L);
      lp[a] += (2.0L * n + 1.0L) / 2.0L * Pn;
      if (n >= 1) {
        const long double p2 = ((2.0L * n + 1.0L) * x[a] * p1 - n * p0) / (n + 1.0L);
        p0 = p1;
        p1 = p2;
      }
    }
  }
  Lift.assign((size_t)Np * 6, 0.0);
  for (int f = 0; f < 6; ++f)
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int a = (f == 1 || f == 3) ? i : (f == 0 || f == 2) ? j : k;
          const bool plus = (f == 1 || f == 2 || f == 5);
          Lift[i + j * Nq + (size_t)k * Nq * Nq + (size_t)f * Np] = (double)(plus ? lp[a] : lm[a]);
        }
  return DGK_OK;
}

int write_operators(const char *dir, int p) {
  if (!dir) return set_error(DGK_ERR_ARG, "operator directory is NULL");
  std::vector<double> x, D, L;
  DGK_TRY(make_operators(p, x, D, L));
  const std::string path = std::string(dir) + "/p" + std::to_string(p) + ".dat";
  FILE *fp = std::fopen(path.c_str(), "w");
  if (!fp) return set_error(DGK_ERR_IO, "cannot write operator file " + path);
  // one record per line, Fortran ES25.16E3 style: " -1.0000000000000000E+000"
  auto record = [&](const std::vector<double> &v) {
    for (double a : v) {
      char m[40];
      std::snprintf(m, sizeof m, "%.16E", a);  // d.ddddE+xx
      std::string s(m);
      const size_t e = s.find('E');
      const char sign = s[e + 1];
      const int ex = std::atoi(s.c_str() + e + 2);
      char out[48];
      std::snprintf(out, sizeof out, "%sE%c%03d", s.substr(0, e).c_str(), sign, ex);
      std::fprintf(fp, "%25s", out);
    }
    std::fputc('\n', fp);
  };
  record(x);
  record(D);
  record(L);
  const bool ok = std::ferror(fp) == 0;
  std::fclose(fp);
  if (!ok) return set_error(DGK_ERR_IO, "write error on " + path);
  return DGK_OK;
}

}  // namespace dgk

extern "C" {

int dgk_make_operators(int p, double *x1D, double *D1D, double *Lift_mat) {
  dgk::clear_error();
  if (!x1D || !D1D || !Lift_mat) return dgk::set_error(DGK_ERR_ARG, "NULL output array");
  std::vector<double> x, D, L;
  DGK_TRY(dgk::make_operators(p, x, D, L));
  std::copy(x.begin(), x.end(), x1D);
  std::copy(D.begin(), D.end(), D1D);
  std::copy(L.begin(), L.end(), Lift_mat);
  return DGK_OK;
}

int dgk_write_operators(const char *dir, int p) {
  dgk::clear_error();
  return dgk::write_operators(dir, p);
}

}  // extern "C"
