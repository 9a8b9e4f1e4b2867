// Host check of the wave-FFT index arithmetic: the stages of WFFT (64 lanes)
// and WFFT32 (32-lane halves) run lane by lane on the CPU, in the order the
// device drivers run them (all reads of an exchange before its writes), and
// are compared with a direct DFT in long double.  Built and run by
// tests/test_wfft_host.py; prints the largest relative error and exits 1
// above 1e-13.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "../shear_nnf_amd/csrc/snf_wfft32.h"

static std::vector<cplx> roots(int n) {
  std::vector<cplx> r(n);
  for (int j = 0; j < n; ++j) {
    const long double a = -2.0L * 3.14159265358979323846264338327950288L * j / n;
    r[j] = make_double2((double)cosl(a), (double)sinl(a));
  }
  return r;
}
static std::vector<cplx> dft(const std::vector<cplx>& x, int sign) {
  const int n = (int)x.size();
  std::vector<cplx> y(n);
  for (int k = 0; k < n; ++k) {
    long double re = 0, im = 0;
    for (int j = 0; j < n; ++j) {
      const long double a = sign * 2.0L * 3.14159265358979323846264338327950288L *
                            (long double)((long)j * k % n) / n;
      re += x[j].x * cosl(a) - x[j].y * sinl(a);
      im += x[j].x * sinl(a) + x[j].y * cosl(a);
    }
    y[k] = make_double2((double)re, (double)im);
  }
  return y;
}
static double relerr(const std::vector<cplx>& a, const std::vector<cplx>& b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += (a[i].x - b[i].x) * (a[i].x - b[i].x) + (a[i].y - b[i].y) * (a[i].y - b[i].y);
    den += b[i].x * b[i].x + b[i].y * b[i].y;
  }
  return std::sqrt(num / den);
}

// G lanes, lane layout in -> task layout out (L -> T), then back (T -> L)
template <class W, int G, int R0, int NT, int NB, int SIGN>
static double check(unsigned seed) {
  const int N = G * R0;
  std::vector<cplx> x(N), rt = roots(N), buf(W::LDS + 64);
  srand(seed);
  for (auto& v : x) v = make_double2(rand() / (double)RAND_MAX - 0.5, rand() / (double)RAND_MAX - 0.5);
  // L -> T
  std::vector<cplx> a((size_t)G * R0), d((size_t)NB * 8), out((size_t)NT * 8);
  for (int l = 0; l < G; ++l) {
    for (int m = 0; m < R0; ++m) a[l * R0 + m] = x[l + G * m];
    W::l2t_a(l, &a[l * R0], buf.data(), rt.data());
  }
  for (int t = 0; t < NB; ++t) W::l2t_b_read(t, buf.data(), &d[t * 8]);
  for (int t = 0; t < NB; ++t) W::l2t_b_write(t, &d[t * 8], buf.data(), rt.data());
  for (int t = 0; t < NT; ++t) W::l2t_c(t, buf.data(), &out[t * 8]);
  std::vector<cplx> y(N), ref = dft(x, SIGN);
  for (int t = 0; t < NT; ++t)
    for (int q = 0; q < 8; ++q) y[t + NT * q] = out[t * 8 + q];
  double e = relerr(y, ref);
  // T -> L of the same input laid out in tasks
  std::vector<cplx> v((size_t)NT * 8), u((size_t)NB * 8), b((size_t)G * R0), z(N);
  for (int t = 0; t < NT; ++t) {
    for (int q = 0; q < 8; ++q) v[t * 8 + q] = x[t + NT * q];
    W::t2l_a(t, &v[t * 8], buf.data(), rt.data());
  }
  for (int t = 0; t < NB; ++t) W::t2l_b_read(t, buf.data(), &u[t * 8]);
  for (int t = 0; t < NB; ++t) W::t2l_b_write(t, &u[t * 8], buf.data());
  for (int l = 0; l < G; ++l) {
    W::t2l_c(l, buf.data(), &b[l * R0], rt.data());
    for (int m = 0; m < R0; ++m) z[l + G * m] = b[l * R0 + m];
  }
  const double e2 = relerr(z, ref);
  return e > e2 ? e : e2;
}

int main() {
  double e = 0;
  auto upd = [&](double v, const char* what) {
    printf("%s %.3e\n", what, v);
    if (v > e) e = v;
  };
  upd(check<WFFT32<2, -1>, 32, 2, 8, 16, -1>(1), "WFFT32 N=64 forward");
  upd(check<WFFT32<2, 1>, 32, 2, 8, 16, 1>(2), "WFFT32 N=64 inverse");
  upd(check<WFFT32<3, -1>, 32, 3, 12, 24, -1>(3), "WFFT32 N=96 forward");
  upd(check<WFFT32<3, 1>, 32, 3, 12, 24, 1>(4), "WFFT32 N=96 inverse");
  upd(check<WFFT32<4, -1>, 32, 4, 16, 32, -1>(5), "WFFT32 N=128 forward");
  upd(check<WFFT<3, -1>, 64, 3, 24, 24, -1>(6), "WFFT N=192 forward");
  upd(check<WFFT<6, 1>, 64, 6, 48, 48, 1>(7), "WFFT N=384 inverse");
  printf("max %.3e\n", e);
  return e < 1e-13 ? 0 : 1;
}
