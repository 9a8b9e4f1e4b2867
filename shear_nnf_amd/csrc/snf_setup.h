// Host-side setup: compact-scheme coefficient design and the untwisted
// base factorisation (run once per handle, binary64).
//
// Restated from diffy_shear_hepta.f90: prederiv1 (68-185), prederiv2
// (187-296), modChol7cyclic_cmplx with twist 1 (578-755); the Gauss-Jordan
// solve of the reference is replaced by Gaussian elimination with partial
// pivoting on the same 6x6 / 7x7 systems.
#pragma once
#include <cmath>
#include <vector>

namespace snf_setup {

inline bool gauss_solve(std::vector<double>& A, std::vector<double>& b, int n) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
    if (A[piv * n + c] == 0.0) return false;
    if (piv != c) {
      for (int q = 0; q < n; ++q) std::swap(A[c * n + q], A[piv * n + q]);
      std::swap(b[c], b[piv]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      for (int q = c; q < n; ++q) A[r * n + q] -= f * A[c * n + q];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int q = r + 1; q < n; ++q) s -= A[r * n + q] * b[q];
    b[r] = s / A[r * n + r];
  }
  return true;
}

// first derivative: Taylor orders 1,3,5 + kappa = (0.5, 0.9, 0.7) pi
inline void design_first(double pr[7], double rh[7]) {
  const int n = 3;
  const double pi = 4.0 * std::atan(1.0);
  const double kap[3] = {0.5 * pi, 0.9 * pi, 0.5 * (0.5 + 0.9) * pi};
  std::vector<double> A(36, 0.0), b(6, 0.0);
  const int orders[3] = {1, 3, 5};
  for (int r = 0; r < 3; ++r)
    for (int j = 1; j <= n; ++j) {
      A[r * 6 + (j - 1)] = orders[r] * std::pow((double)j, orders[r] - 1);
      A[r * 6 + n + (j - 1)] = -std::pow((double)j, orders[r]);
    }
  b[0] = -0.5;
  for (int q = 0; q < 3; ++q) {
    for (int j = 1; j <= n; ++j) {
      A[(3 + q) * 6 + (j - 1)] = kap[q] * std::cos(kap[q] * j);
      A[(3 + q) * 6 + n + (j - 1)] = -std::sin(kap[q] * j);
    }
    b[3 + q] = -0.5 * kap[q];
  }
  gauss_solve(A, b, 6);
  pr[3] = 1.0;
  rh[3] = 0.0;
  for (int j = 1; j <= 3; ++j) {
    pr[3 + j] = pr[3 - j] = b[j - 1];
    rh[3 + j] = b[n + j - 1];
    rh[3 - j] = -b[n + j - 1];
  }
}

// second derivative: consistency, Taylor orders 2,4,6 + kappa = (0.9, 0.5, 0.7) pi
inline void design_second(double pr[7], double rh[7]) {
  const int n = 3;
  const double pi = 4.0 * std::atan(1.0);
  const double kap[3] = {0.9 * pi, 0.5 * pi, 0.5 * (0.9 + 0.5) * pi};
  std::vector<double> A(49, 0.0), b(7, 0.0);
  for (int q = n + 1; q < 7; ++q) A[q] = 1.0;
  A[n] = 0.5;
  const int orders[3] = {2, 4, 6};
  for (int r = 0; r < 3; ++r)
    for (int j = 1; j <= n; ++j) {
      A[(1 + r) * 7 + (j - 1)] = orders[r] * (orders[r] - 1) * std::pow((double)j, orders[r] - 2);
      A[(1 + r) * 7 + n + j] = -std::pow((double)j, orders[r]);
    }
  b[1] = -1.0;
  for (int q = 0; q < 3; ++q) {
    for (int j = 1; j <= n; ++j) {
      A[(4 + q) * 7 + (j - 1)] = kap[q] * kap[q] * std::cos(kap[q] * j);
      A[(4 + q) * 7 + n + j] = std::cos(kap[q] * j);
    }
    A[(4 + q) * 7 + n] = 0.5;
    b[4 + q] = -0.5 * kap[q] * kap[q];
  }
  gauss_solve(A, b, 7);
  pr[3] = 1.0;
  rh[3] = b[n];
  for (int j = 1; j <= 3; ++j) {
    pr[3 + j] = pr[3 - j] = b[j - 1];
    rh[3 + j] = rh[3 - j] = b[n + j];
  }
}

// Untwisted L D L^T of the cyclic symmetric hepta matrix with band a[3..6];
// border rows real.  Arrays of length N, zero where a row has no entry.
struct Base {
  std::vector<double> D, Dm1, La, Lb, Lc, Le, Lf, Lg;
};

inline Base base_factor(const double a[7], int N) {
  Base f;
  for (auto* v : {&f.D, &f.Dm1, &f.La, &f.Lb, &f.Lc, &f.Le, &f.Lf, &f.Lg}) v->assign(N, 0.0);
  const double a0 = a[3], a1 = a[4], a2 = a[5], a3 = a[6];
  auto g = [](const std::vector<double>& v, int i) { return i >= 0 ? v[i] : 0.0; };
  const int last = N - 4;
  for (int j = 0; j <= last; ++j) {
    f.D[j] = a0 - (g(f.Lc, j) * g(f.D, j - 3) * g(f.Lc, j) + g(f.Lb, j) * g(f.D, j - 2) * g(f.Lb, j) +
                   g(f.La, j) * g(f.D, j - 1) * g(f.La, j));
    f.Dm1[j] = 1.0 / f.D[j];
    if (j + 1 <= last)
      f.La[j + 1] = f.Dm1[j] * (a1 - (g(f.Lc, j + 1) * g(f.D, j - 2) * g(f.Lb, j) +
                                      g(f.Lb, j + 1) * g(f.D, j - 1) * g(f.La, j)));
    if (j + 2 <= last) f.Lb[j + 2] = f.Dm1[j] * (a2 - g(f.Lc, j + 2) * g(f.D, j - 1) * g(f.La, j));
    if (j + 3 <= last) f.Lc[j + 3] = f.Dm1[j] * a3;
  }
  auto border = [&](std::vector<double>& L, auto src) {
    for (int i = 0; i <= last; ++i)
      L[i] = f.Dm1[i] * (src(i) - g(L, i - 3) * g(f.D, i - 3) * g(f.Lc, i) -
                         g(L, i - 2) * g(f.D, i - 2) * g(f.Lb, i) - g(L, i - 1) * g(f.D, i - 1) * g(f.La, i));
  };
  border(f.Le, [&](int i) {
    double v = 0.0;
    if (i == 0) v += a3;
    if (i == N - 6) v += a3;
    if (i == N - 5) v += a2;
    if (i == N - 4) v += a1;
    return v;
  });
  double s = 0.0;
  for (int k = 0; k <= last; ++k) s += f.Le[k] * f.Le[k] * f.D[k];
  f.D[N - 3] = a0 - s;
  f.Dm1[N - 3] = 1.0 / f.D[N - 3];
  border(f.Lf, [&](int i) {
    double v = 0.0;
    if (i == 0) v += a2;
    if (i == 1) v += a3;
    if (i == N - 5) v += a3;
    if (i == N - 4) v += a2;
    return v;
  });
  s = 0.0;
  for (int k = 0; k <= last; ++k) s += f.Lf[k] * f.Le[k] * f.D[k];
  f.Lf[N - 3] = f.Dm1[N - 3] * (a1 - s);
  s = 0.0;
  for (int k = 0; k <= N - 3; ++k) s += f.Lf[k] * f.Lf[k] * f.D[k];
  f.D[N - 2] = a0 - s;
  f.Dm1[N - 2] = 1.0 / f.D[N - 2];
  border(f.Lg, [&](int i) {
    double v = 0.0;
    if (i == 0) v += a1;
    if (i == 1) v += a2;
    if (i == 2) v += a3;
    if (i == N - 4) v += a3;
    return v;
  });
  s = 0.0;
  for (int k = 0; k <= last; ++k) s += f.Lg[k] * f.D[k] * f.Le[k];
  f.Lg[N - 3] = f.Dm1[N - 3] * (a2 - s);
  s = 0.0;
  for (int k = 0; k <= N - 3; ++k) s += f.Lg[k] * f.D[k] * f.Lf[k];
  f.Lg[N - 2] = f.Dm1[N - 2] * (a1 - s);
  s = 0.0;
  for (int k = 0; k <= N - 2; ++k) s += f.Lg[k] * f.Lg[k] * f.D[k];
  f.D[N - 1] = a0 - s;
  f.Dm1[N - 1] = 1.0 / f.D[N - 1];
  return f;
}

}  // namespace snf_setup
