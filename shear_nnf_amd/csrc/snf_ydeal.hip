// y-dealiasing variants 1 and 2 of the time advance (iydealiasing,
// cross.f90:459-475 and dealiasing_y, operators.f90:352-486).
//
// At the top of every substep the state (omega_y, phi) is taken to the
// moving frame exp(i kx y fac), with fac the fractional bottom offset
// (xwkb/Lx - int(xwkb/Lx)) Lx/Ly shifted into [-Lx/2Ly, Lx/Ly), each line is
// transformed in y (cfty: forward FFT scaled by 1/my), the coefficients in
// the window of its kx index are zeroed, and the line is transformed back
// and taken back to the lab frame.  Variant 1 cuts only the kx = 0 lines
// (the window (ny1, ny2)); variant 2 every line, with windows shifted by
// djdb and djda in proportion to kx and fac.  U and W lose their modes
// k >= my/3 (rfty, real index my23 on).
//
// One wave per line: the line, its coefficients and the root table
// exp(-2 pi i t/my) sit in LDS and each lane forms coefficients (and then
// values) of its own indices as direct sums, so any my is served.  The
// variants are off by default; the cut is a full-spectrum projection whose
// cost (2 my^2 per line and field) matters only when they are selected.
#pragma once
#include "snf_internal.h"

struct YDealiasArgs {
  Geom g;
  cplx *vor, *phi;
  cplx *u00c, *w00c;      // real lines stored with imaginary part 0
  const cplx* twy;        // exp(-2 pi i t/my), t < my
  DevScalars* sc;
  int variant;            // 1 or 2
  int ny1, ny2;           // ny, my - ny with ny = ((2 (my/3) - 1) - 1)/2 (initcr.f90:207-210)
  int kcut;               // my23/2 = my/3: first real-line mode cut
};

// dynamic LDS: x[my], X[my], rt[my]
__global__ void __launch_bounds__(64) k_ydealias(YDealiasArgs a) {
  if (a.sc->halted) return;
  extern __shared__ cplx yd_lds[];
  const Geom& g = a.g;
  const int N = g.my, lane = threadIdx.x;
  cplx* x = yd_lds;
  cplx* X = yd_lds + N;
  cplx* rt = yd_lds + 2 * N;
  for (int t = lane; t < N; t += 64) rt[t] = gld(a.twy + t);

  const double Lx = 2.0 * M_PI / g.alp, Ly = g.Ly;
  const double q = a.sc->xwkb / Lx;
  double fac = (q - trunc(q)) * Lx / Ly;
  if (fac < -Lx / Ly * 0.5) fac = Lx / Ly + fac;

  const bool line00 = (int)blockIdx.x == (int)gridDim.x - 1;   // the U, W block
  const long L = line00 ? 0 : (long)blockIdx.x;
  const int i = (int)(L / g.mz);
  const double kx = g.alp * (double)i;
  const double mx1 = (double)(g.kxn - 1), ny1 = (double)a.ny1, ny2 = (double)a.ny2;
  double djda;
  if (a.variant == 1 || fac <= 0.0)
    djda = ny1 / mx1 * (-fac - Lx / Ly) * Ly / Lx * (double)i;
  else
    djda = ny1 / mx1 * (-fac + Lx / Ly) * Ly / Lx * (double)i;
  const double djdb = ny1 / mx1 * (-fac) * Ly / Lx * (double)i;
  auto cut = [&](int j) -> bool {
    const double dj = (double)j;
    if (line00) return (j < N - j ? j : N - j) >= a.kcut;
    if (a.variant == 1) return dj > ny1 + djdb && dj < ny2 + djdb;
    return (dj >= ny1 + djdb && dj <= ny2 + djdb) || (dj > ny1 + djda && dj < ny2 + djda);
  };
  const double invN = 1.0 / (double)N;

  for (int f = 0; f < 2; ++f) {
    cplx* p;
    long stride;
    if (line00) {
      p = f == 0 ? a.u00c : a.w00c;
      stride = 1;
    } else {
      p = (f == 0 ? a.vor : a.phi) + sline(g, i, (int)(L - (long)i * g.mz));
      stride = SNF_CW;
    }
    __syncthreads();
    // moving frame exp(i kx y fac) (operators.f90:86-145); identity at kx = 0
    for (int j = lane; j < N; j += 64) {
      cplx v = gld(p + (long)j * stride);
      if (!line00 && i > 0) {
        double sn, cs;
        sincos(kx * (j * g.dy - 0.5 * Ly) * fac, &sn, &cs);
        v = cmul(v, mkc(cs, sn));
      }
      x[j] = v;
    }
    __syncthreads();
    // forward transform scaled by 1/N, then the cut
    for (int m = lane; m < N; m += 64) {
      double re = 0.0, im = 0.0;
      int t = 0;
      for (int j = 0; j < N; ++j) {
        const cplx xv = x[j], w = rt[t];
        re += xv.x * w.x - xv.y * w.y;
        im += xv.x * w.y + xv.y * w.x;
        t += m;
        if (t >= N) t -= N;
      }
      X[m] = cut(m) ? mkc(0.0, 0.0) : mkc(re * invN, im * invN);
    }
    __syncthreads();
    // backward transform (unscaled), back to the lab frame
    for (int j = lane; j < N; j += 64) {
      double re = 0.0, im = 0.0;
      int t = 0;
      for (int m = 0; m < N; ++m) {
        const cplx c = X[m], w = rt[t];   // conj(w) = exp(+2 pi i jm/N)
        re += c.x * w.x + c.y * w.y;
        im += c.y * w.x - c.x * w.y;
        t += j;
        if (t >= N) t -= N;
      }
      cplx v = mkc(re, im);
      if (line00) {
        v.y = 0.0;   // the real-to-complex pair returns a real line
      } else if (i > 0) {
        double sn, cs;
        sincos(kx * (j * g.dy - 0.5 * Ly) * (-fac), &sn, &cs);
        v = cmul(v, mkc(cs, sn));
      }
      gst(p + (long)j * stride, v);
    }
  }
}
