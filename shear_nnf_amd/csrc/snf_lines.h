// Batched y-line solvers: one thread owns one (kx,kz) line of my points.
//
// Storage is [y][kx][kz] complex, so element j of line L = i*mz + k sits at
// j*P + L (P = kxn*mz): consecutive threads read consecutive addresses at
// every j (coalesced), and the serial dependence runs along j inside a
// thread.  The forward substitution writes its intermediate y(j) to a
// scratch field that the backward substitution reads back in reverse.
//
// There are few lines (kxn*mz = 32640 at 384^3, about one wave per SIMD) and
// the sweeps are serial, so every stream a sweep reads (input field, y
// scratch, Helmholtz factor, fields the backward step combines) is
// prefetched D rows ahead through a register ring: row j+D is requested as
// soon as row j is consumed.  All loads and stores go through the global
// address space (gld/gst) so that the waits for them count outstanding
// requests instead of draining the queue, and the factor rows, which are the
// same for every line, are read through the scalar cache (cld).
//
// The interior rows j <= N-7 use the untwisted border rows scaled by the
// twist (Le(j) = Ler(j)*shp, ...); rows N-6..N-4 and the three border rows
// use the re-factored tail.  Those rows are peeled out of the main loop so
// that the loop body has no per-row selection.
//
// Algorithms restated from diffy_shear_hepta.f90 (reference):
//   RHS stencil with shear wrap     derivyc 358-383 / lapcdy 487-514
//   L D L^H of the twisted cyclic   modChol7cyclic_cmplx 578-755
//   twisted re-factor of borders    derivyc 389-406 + modChol_addshear 760-870
//   substitution                    solve7cyclic 900-954
#pragma once
#include "snf_internal.h"

// read-only, wave-uniform data (factor rows): scalar loads
#define SNF_AS4 __attribute__((address_space(4)))
__device__ __forceinline__ double cld(const double* p) { return *(const SNF_AS4 double*)p; }

// ---------------------------------------------------------------------------
// twisted derivative operator seen by one line
// ---------------------------------------------------------------------------
struct DOp {
  BaseFactor b;
  cplx shp;
  // re-factored tail (see TwistTail): Le rows N-6..N-4, Lf rows N-5..N-3,
  // Lg rows N-4..N-2, 1/D rows N-3..N-1
  cplx le0, le1, le2, lf0, lf1, lf2, lg0, lg1, lg2;
  double dm0, dm1, dm2;
  int N;
};

__device__ __forceinline__ DOp make_dop(const BaseFactor& b, const TwistTail* t, cplx shp, int N) {
  DOp o;
  o.b = b;
  o.shp = shp;
  o.le0 = gld(&t->Le[0]); o.le1 = gld(&t->Le[1]); o.le2 = gld(&t->Le[2]);
  o.lf0 = gld(&t->Lf[0]); o.lf1 = gld(&t->Lf[1]); o.lf2 = gld(&t->Lf[2]);
  o.lg0 = gld(&t->Lg[0]); o.lg1 = gld(&t->Lg[1]); o.lg2 = gld(&t->Lg[2]);
  o.dm0 = gldd(&t->Dm1[0]); o.dm1 = gldd(&t->Dm1[1]); o.dm2 = gldd(&t->Dm1[2]);
  o.N = N;
  return o;
}

// tail held by value (untwisted operators of the real 00 lines)
__device__ __forceinline__ DOp make_dop(const BaseFactor& b, const TwistTail& t, cplx shp, int N) {
  DOp o;
  o.b = b;
  o.shp = shp;
  o.le0 = t.Le[0]; o.le1 = t.Le[1]; o.le2 = t.Le[2];
  o.lf0 = t.Lf[0]; o.lf1 = t.Lf[1]; o.lf2 = t.Lf[2];
  o.lg0 = t.Lg[0]; o.lg1 = t.Lg[1]; o.lg2 = t.Lg[2];
  o.dm0 = t.Dm1[0]; o.dm1 = t.Dm1[1]; o.dm2 = t.Dm1[2];
  o.N = N;
  return o;
}

// border-row coefficients of an interior row (j <= N-7)
__device__ __forceinline__ cplx dop_e(const DOp& o, int j) { return crmul(cld(o.b.Ler + j), o.shp); }
__device__ __forceinline__ cplx dop_f(const DOp& o, int j) { return crmul(cld(o.b.Lfr + j), o.shp); }
__device__ __forceinline__ cplx dop_g(const DOp& o, int j) { return crmul(cld(o.b.Lgr + j), o.shp); }

// Le, Lf, Lg of forward row j given as T = 0 (interior), 1 (N-6), 2 (N-5), 3 (N-4)
template <int T>
__device__ __forceinline__ void dop_row(const DOp& o, int j, cplx& le, cplx& lf, cplx& lg) {
  if (T == 0) { le = dop_e(o, j); lf = dop_f(o, j); lg = dop_g(o, j); }
  if (T == 1) { le = o.le0; lf = dop_f(o, j); lg = dop_g(o, j); }
  if (T == 2) { le = o.le1; lf = o.lf0; lg = dop_g(o, j); }
  if (T == 3) { le = o.le2; lf = o.lf1; lg = o.lg0; }
}

// 7-point stencil sum_m c(m) w(m), m=-3..3, left to right (centre skipped
// for the first derivative whose centre coefficient is zero).
template <bool FIRST>
__device__ __forceinline__ cplx stencil7(const double* c, const cplx* w) {
  cplx r = mkc(0.0, 0.0);
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    if (FIRST && m == 3) continue;
    r.x += c[m] * w[m].x;
    r.y += c[m] * w[m].y;
  }
  return r;
}

// Window setup: w[0..2] = u(-3..-1) = u(N-3..N-1)*shm, w[3..5] = u(0..2);
// bot[0..2] = u(N-3..N-1); top[0..2] = u(N..N+2) = u(0..2)*shp.
__device__ __forceinline__ void window_init(const cplx* in, long P, int N, cplx shp, cplx shm,
                                            cplx* w, cplx* bot, cplx* top) {
  bot[0] = gld(in + (long)(N - 3) * P);
  bot[1] = gld(in + (long)(N - 2) * P);
  bot[2] = gld(in + (long)(N - 1) * P);
  w[3] = gld(in);
  w[4] = gld(in + P);
  w[5] = gld(in + 2 * P);
  w[0] = cmul(bot[0], shm);
  w[1] = cmul(bot[1], shm);
  w[2] = cmul(bot[2], shm);
  top[0] = cmul(w[3], shp);
  top[1] = cmul(w[4], shp);
  top[2] = cmul(w[5], shp);
}

__device__ __forceinline__ void window_shift(cplx* w) {
#pragma unroll
  for (int m = 0; m < 6; ++m) w[m] = w[m + 1];
}

// forward substitution of one interior row (solve7cyclic 911-923; the
// border sums of 924-944 are accumulated by the caller)
__device__ __forceinline__ cplx fwd_row(cplx b, double La, double Lb, double Lc, cplx& y1,
                                        cplx& y2, cplx& y3) {
  cplx y;
  y.x = b.x - (Lc * y3.x + Lb * y2.x + La * y1.x);
  y.y = b.y - (Lc * y3.y + Lb * y2.y + La * y1.y);
  y3 = y2;
  y2 = y1;
  y1 = y;
  return y;
}

// backward substitution of one interior row (solve7cyclic 946-952)
__device__ __forceinline__ cplx bwd_row(cplx y, double dm, double La, double Lb, double Lc,
                                        cplx ce, cplx cf, cplx cg, cplx xN3, cplx xN2, cplx xN1,
                                        cplx& x1, cplx& x2, cplx& x3) {
  cplx x = crmul(dm, y);
  x.x = x.x - La * x1.x - Lb * x2.x - Lc * x3.x;
  x.y = x.y - La * x1.y - Lb * x2.y - Lc * x3.y;
  x = csub(csub(csub(x, cmul(ce, xN3)), cmul(cf, xN2)), cmul(cg, xN1));
  x3 = x2;
  x2 = x1;
  x1 = x;
  return x;
}

// ---------------------------------------------------------------------------
// derivative solve (twisted derivative factor shared by NF fields)
//   forward:  y = L^-1 (R u)   -> ys
//   backward: x = L^-H D^-1 y  -> sink
// Requires N >= D + 7.
// ---------------------------------------------------------------------------
template <int NF, bool FIRST>
struct FwdState {
  cplx w[NF][7], bot[NF][3], top[NF][3];
  cplx y1[NF], y2[NF], y3[NF], E[NF], F[NF], G[NF];
};

template <int NF, bool FIRST, int T>
__device__ __forceinline__ void deriv_fwd_row(const DOp& o, const double* c, int j,
                                              FwdState<NF, FIRST>& s, const cplx* w6,
                                              cplx* const* ys, long py) {
  const double La = cld(o.b.La + j), Lb = cld(o.b.Lb + j), Lc = cld(o.b.Lc + j);
  cplx le, lf, lg;
  dop_row<T>(o, j, le, lf, lg);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    s.w[f][6] = w6[f];
    const cplx y = fwd_row(stencil7<FIRST>(c, s.w[f]), La, Lb, Lc, s.y1[f], s.y2[f], s.y3[f]);
    s.E[f] = cadd(s.E[f], cmul(le, y));
    s.F[f] = cadd(s.F[f], cmul(lf, y));
    s.G[f] = cadd(s.G[f], cmul(lg, y));
    gst(ys[f] + (long)j * py, y);
    window_shift(s.w[f]);
  }
}

// border rows N-3, N-2, N-1 of the forward sweep (wrap with shp at the top)
template <int NF, bool FIRST>
__device__ __forceinline__ void border_rows(const DOp& o, const double* c, FwdState<NF, FIRST>& s,
                                            cplx* const* ys, long py) {
  const int N = o.N;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    s.w[f][6] = s.top[f][0];
    const cplx yN3 = csub(stencil7<FIRST>(c, s.w[f]), s.E[f]);
    s.F[f] = cadd(s.F[f], cmul(o.lf2, yN3));
    s.G[f] = cadd(s.G[f], cmul(o.lg1, yN3));
    window_shift(s.w[f]);
    s.w[f][6] = s.top[f][1];
    const cplx yN2 = csub(stencil7<FIRST>(c, s.w[f]), s.F[f]);
    s.G[f] = cadd(s.G[f], cmul(o.lg2, yN2));
    window_shift(s.w[f]);
    s.w[f][6] = s.top[f][2];
    const cplx yN1 = csub(stencil7<FIRST>(c, s.w[f]), s.G[f]);
    gst(ys[f] + (long)(N - 3) * py, yN3);
    gst(ys[f] + (long)(N - 2) * py, yN2);
    gst(ys[f] + (long)(N - 1) * py, yN1);
  }
}

template <int NF, bool FIRST, int D>
__device__ __forceinline__ void deriv_forward(const DOp& o, const double* c,
                                              const cplx* const* in, long pin,
                                              cplx* const* ys, long py, cplx shm) {
  const int N = o.N;
  const int mlast = N - 7;   // last interior row
  FwdState<NF, FIRST> s;
  cplx pf[NF][D];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    window_init(in[f], pin, N, o.shp, shm, s.w[f], s.bot[f], s.top[f]);
    s.y1[f] = s.y2[f] = s.y3[f] = s.E[f] = s.F[f] = s.G[f] = mkc(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < D; ++q) pf[f][q] = gld(in[f] + (long)(3 + q) * pin);
  }
  for (int jb = 0; jb <= mlast; jb += D) {
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const int j = jb + q;
      if (j <= mlast) {
        cplx w6[NF];
        const int rn = j + 3 + D;   // row refilled into this ring slot
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          w6[f] = pf[f][q];
          if (rn <= N - 4) pf[f][q] = gld(in[f] + (long)rn * pin);
        }
        deriv_fwd_row<NF, FIRST, 0>(o, c, j, s, w6, ys, py);
      }
    }
  }
  cplx w6[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) w6[f] = s.bot[f][0];
  deriv_fwd_row<NF, FIRST, 1>(o, c, N - 6, s, w6, ys, py);
#pragma unroll
  for (int f = 0; f < NF; ++f) w6[f] = s.bot[f][1];
  deriv_fwd_row<NF, FIRST, 2>(o, c, N - 5, s, w6, ys, py);
#pragma unroll
  for (int f = 0; f < NF; ++f) w6[f] = s.bot[f][2];
  deriv_fwd_row<NF, FIRST, 3>(o, c, N - 4, s, w6, ys, py);
  border_rows<NF, FIRST>(o, c, s, ys, py);
}

// Sinks receive the solution row by row, bottom to top.  A sink declares
// the per-row inputs it combines with x (Sink::In), fetched D rows ahead.
template <int NF>
struct BwdState {
  cplx xN1[NF], xN2[NF], xN3[NF], x1[NF], x2[NF], x3[NF];
};

template <int NF, class Sink>
__device__ __forceinline__ void bwd_apply(int i, double dm, double La, double Lb, double Lc,
                                          cplx ce, cplx cf, cplx cg, const cplx* y,
                                          const typename Sink::In& in, BwdState<NF>& s,
                                          Sink& sink) {
  cplx xo[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f)
    xo[f] = bwd_row(y[f], dm, La, Lb, Lc, ce, cf, cg, s.xN3[f], s.xN2[f], s.xN1[f], s.x1[f],
                    s.x2[f], s.x3[f]);
  sink.apply(i, in, xo);
}

template <int NF, int D, class Sink>
__device__ __forceinline__ void deriv_backward(const DOp& o, cplx* const* ys, long py,
                                               Sink& sink) {
  const int N = o.N;
  const int mtop = N - 7;   // first interior row of the backward sweep
  BwdState<NF> s;
  cplx pyr[NF][D];
  typename Sink::In pin_[D];
  // ring: rows N-7-q
#pragma unroll
  for (int q = 0; q < D; ++q) {
    const int r = mtop - q;
#pragma unroll
    for (int f = 0; f < NF; ++f) pyr[f][q] = gld(ys[f] + (long)r * py);
    pin_[q] = sink.load(r);
  }
  // border rows N-1, N-2, N-3
  {
    const typename Sink::In t1 = sink.load(N - 1), t2 = sink.load(N - 2), t3 = sink.load(N - 3);
    const cplx cg2 = cconj(o.lg2), cf3 = cconj(o.lf2), cg3 = cconj(o.lg1);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const cplx yN1 = gld(ys[f] + (long)(N - 1) * py);
      const cplx yN2 = gld(ys[f] + (long)(N - 2) * py);
      const cplx yN3 = gld(ys[f] + (long)(N - 3) * py);
      s.xN1[f] = crmul(o.dm2, yN1);
      s.xN2[f] = csub(crmul(o.dm1, yN2), cmul(cg2, s.xN1[f]));
      s.xN3[f] = csub(csub(crmul(o.dm0, yN3), cmul(cf3, s.xN2[f])), cmul(cg3, s.xN1[f]));
      s.x1[f] = s.xN3[f];
      s.x2[f] = s.xN2[f];
      s.x3[f] = s.xN1[f];
    }
    sink.apply(N - 1, t1, s.xN1);
    sink.apply(N - 2, t2, s.xN2);
    sink.apply(N - 3, t3, s.xN3);
  }
  // tail rows N-4, N-5, N-6
  {
    cplx y[NF];
    int i = N - 4;
    typename Sink::In in = sink.load(i);
#pragma unroll
    for (int f = 0; f < NF; ++f) y[f] = gld(ys[f] + (long)i * py);
    bwd_apply<NF>(i, cld(o.b.Dm1 + i), cld(o.b.La + i + 1), cld(o.b.Lb + i + 2),
                  cld(o.b.Lc + i + 3), cconj(o.le2), cconj(o.lf1), cconj(o.lg0), y, in, s, sink);
    i = N - 5;
    in = sink.load(i);
#pragma unroll
    for (int f = 0; f < NF; ++f) y[f] = gld(ys[f] + (long)i * py);
    bwd_apply<NF>(i, cld(o.b.Dm1 + i), cld(o.b.La + i + 1), cld(o.b.Lb + i + 2),
                  cld(o.b.Lc + i + 3), cconj(o.le1), cconj(o.lf0), cconj(dop_g(o, i)), y, in, s,
                  sink);
    i = N - 6;
    in = sink.load(i);
#pragma unroll
    for (int f = 0; f < NF; ++f) y[f] = gld(ys[f] + (long)i * py);
    bwd_apply<NF>(i, cld(o.b.Dm1 + i), cld(o.b.La + i + 1), cld(o.b.Lb + i + 2),
                  cld(o.b.Lc + i + 3), cconj(o.le0), cconj(dop_f(o, i)), cconj(dop_g(o, i)), y,
                  in, s, sink);
  }
  for (int ib = mtop; ib >= 0; ib -= D) {
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const int i = ib - q;
      if (i >= 0) {
        const typename Sink::In in = pin_[q];
        cplx y[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) y[f] = pyr[f][q];
        const int rn = i - D;
        if (rn >= 0) {
#pragma unroll
          for (int f = 0; f < NF; ++f) pyr[f][q] = gld(ys[f] + (long)rn * py);
          pin_[q] = sink.load(rn);
        }
        bwd_apply<NF>(i, cld(o.b.Dm1 + i), cld(o.b.La + i + 1), cld(o.b.Lb + i + 2),
                      cld(o.b.Lc + i + 3), cconj(dop_e(o, i)), cconj(dop_f(o, i)),
                      cconj(dop_g(o, i)), y, in, s, sink);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Helmholtz solve (d_yy - k2) p = u with on-the-fly factorisation of
// rh2 - k2 pr2 (lapcdy); the per-row factor needed by the backward sweep is
// stored next to y in scratch arrays hr (4 reals/row) and hc (3 complex/row).
// ---------------------------------------------------------------------------
struct HScratch {
  double* hr;  // [4][N][P]: Dm1(j), La(j+1), Lb(j+2), Lc(j+3)
  cplx* hc;    // [3][N][P]: Le(j), Lf(j), Lg(j)
};

struct HTail {      // border values kept in registers between the sweeps
  double Dm1[3];    // rows N-3, N-2, N-1
  cplx LfN3, LgN3, LgN2;
};

// running state of the on-the-fly factorisation
struct HFac {
  double a0, a1, a2, a3;                  // lap matrix diagonals 0..3
  double D1, D2, D3;                      // D(j-1), D(j-2), D(j-3)
  double La_j, Lb_j, Lb_j1, Lc_j, Lc_j1, Lc_j2;
  cplx e1, e2, e3, f1, f2, f3, g1, g2, g3;
  double See, Sff, Sgg;
  cplx Sfe, Sge, Sgf;
};

// Factor row j.  T = 0: interior row j <= N-7 (corner entries for j < 3);
// T = 1, 2, 3: rows N-6, N-5, N-4.  Returns the border-row coefficients and
// writes the backward-sweep row to the scratch.
template <int T>
__device__ __forceinline__ void helm_factor_row(HFac& h, int j, int /*N*/, cplx shp, HScratch hs,
                                                long NP, long o, cplx& le, cplx& lf, cplx& lg,
                                                double& La_out, double& Lb_out, double& Lc_out) {
  const double Dj = h.a0 - (h.Lc_j * h.D3 * h.Lc_j + h.Lb_j * h.D2 * h.Lb_j + h.La_j * h.D1 * h.La_j);
  const double dm = 1.0 / Dj;
  // next-row interior entries (zero past the interior block)
  const double La_n = (T <= 2) ? dm * (h.a1 - (h.Lc_j1 * h.D2 * h.Lb_j + h.Lb_j1 * h.D1 * h.La_j)) : 0.0;
  const double Lb_n = (T <= 1) ? dm * (h.a2 - h.Lc_j2 * h.D1 * h.La_j) : 0.0;
  const double Lc_n = (T == 0) ? dm * h.a3 : 0.0;
  // border rows (entries A(row, j) of the cyclic matrix)
  cplx se = mkc(0, 0), sf = se, sg = se;
  if (T == 0) {
    if (j == 0) { se = crmul(h.a3, shp); sf = crmul(h.a2, shp); sg = crmul(h.a1, shp); }
    if (j == 1) { sf = crmul(h.a3, shp); sg = crmul(h.a2, shp); }
    if (j == 2) { sg = crmul(h.a3, shp); }
  }
  if (T == 1) se.x += h.a3;
  if (T == 2) { se.x += h.a2; sf.x += h.a3; }
  if (T == 3) { se.x += h.a1; sf.x += h.a2; sg.x += h.a3; }
  le.x = dm * (se.x - h.e3.x * h.D3 * h.Lc_j - h.e2.x * h.D2 * h.Lb_j - h.e1.x * h.D1 * h.La_j);
  le.y = dm * (se.y - h.e3.y * h.D3 * h.Lc_j - h.e2.y * h.D2 * h.Lb_j - h.e1.y * h.D1 * h.La_j);
  lf.x = dm * (sf.x - h.f3.x * h.D3 * h.Lc_j - h.f2.x * h.D2 * h.Lb_j - h.f1.x * h.D1 * h.La_j);
  lf.y = dm * (sf.y - h.f3.y * h.D3 * h.Lc_j - h.f2.y * h.D2 * h.Lb_j - h.f1.y * h.D1 * h.La_j);
  lg.x = dm * (sg.x - h.g3.x * h.D3 * h.Lc_j - h.g2.x * h.D2 * h.Lb_j - h.g1.x * h.D1 * h.La_j);
  lg.y = dm * (sg.y - h.g3.y * h.D3 * h.Lc_j - h.g2.y * h.D2 * h.Lb_j - h.g1.y * h.D1 * h.La_j);
  h.See += cabs2(le) * Dj;
  h.Sfe = cadd(h.Sfe, crmul(Dj, cmulc(lf, le)));
  h.Sff += cabs2(lf) * Dj;
  h.Sge = cadd(h.Sge, cmulc(crmul(Dj, lg), le));
  h.Sgf = cadd(h.Sgf, cmulc(crmul(Dj, lg), lf));
  h.Sgg += cabs2(lg) * Dj;
  gstd(hs.hr + o, dm);
  gstd(hs.hr + NP + o, La_n);
  gstd(hs.hr + 2 * NP + o, Lb_n);
  gstd(hs.hr + 3 * NP + o, Lc_n);
  gst(hs.hc + o, le);
  gst(hs.hc + NP + o, lf);
  gst(hs.hc + 2 * NP + o, lg);
  La_out = h.La_j;
  Lb_out = h.Lb_j;
  Lc_out = h.Lc_j;
  h.D3 = h.D2; h.D2 = h.D1; h.D1 = Dj;
  h.La_j = La_n;
  h.Lb_j = h.Lb_j1; h.Lb_j1 = Lb_n;
  h.Lc_j = h.Lc_j1; h.Lc_j1 = h.Lc_j2; h.Lc_j2 = Lc_n;
  h.e3 = h.e2; h.e2 = h.e1; h.e1 = le;
  h.f3 = h.f2; h.f2 = h.f1; h.f1 = lf;
  h.g3 = h.g2; h.g2 = h.g1; h.g1 = lg;
}

template <int NF, int T>
__device__ __forceinline__ void helm_fwd_row(HFac& h, const double* pr, int j, int N, cplx shp,
                                             HScratch hs, long NP, long P,
                                             FwdState<NF, false>& s, const cplx* w6,
                                             cplx* const* ys) {
  cplx le, lf, lg;
  double La, Lb, Lc;
  const long o = (long)j * P;
  helm_factor_row<T>(h, j, N, shp, hs, NP, o, le, lf, lg, La, Lb, Lc);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    s.w[f][6] = w6[f];
    const cplx y = fwd_row(stencil7<false>(pr, s.w[f]), La, Lb, Lc, s.y1[f], s.y2[f], s.y3[f]);
    s.E[f] = cadd(s.E[f], cmul(le, y));
    s.F[f] = cadd(s.F[f], cmul(lf, y));
    s.G[f] = cadd(s.G[f], cmul(lg, y));
    gst(ys[f] + o, y);
    window_shift(s.w[f]);
  }
}

template <int NF, int D>
__device__ __forceinline__ void helm_forward(const double* a /*7: lap matrix*/, const double* pr,
                                             int N, cplx shp, cplx shm,
                                             const cplx* const* in, long pin,
                                             cplx* const* ys, long P, HScratch hs,
                                             HTail& tail) {  // P: stride of ys and hs
  HFac h;
  h.a0 = a[3]; h.a1 = a[4]; h.a2 = a[5]; h.a3 = a[6];
  h.D1 = h.D2 = h.D3 = 0.0;
  h.La_j = h.Lb_j = h.Lb_j1 = h.Lc_j = h.Lc_j1 = h.Lc_j2 = 0.0;
  h.e1 = h.e2 = h.e3 = h.f1 = h.f2 = h.f3 = h.g1 = h.g2 = h.g3 = mkc(0, 0);
  h.See = h.Sff = h.Sgg = 0.0;
  h.Sfe = h.Sge = h.Sgf = mkc(0, 0);
  FwdState<NF, false> s;
  cplx pf[NF][D];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    window_init(in[f], pin, N, shp, shm, s.w[f], s.bot[f], s.top[f]);
    s.y1[f] = s.y2[f] = s.y3[f] = s.E[f] = s.F[f] = s.G[f] = mkc(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < D; ++q) pf[f][q] = gld(in[f] + (long)(3 + q) * pin);
  }
  const int mlast = N - 7;
  const long NP = (long)N * P;
  for (int jb = 0; jb <= mlast; jb += D) {
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const int j = jb + q;
      if (j <= mlast) {
        cplx w6[NF];
        const int rn = j + 3 + D;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          w6[f] = pf[f][q];
          if (rn <= N - 4) pf[f][q] = gld(in[f] + (long)rn * pin);
        }
        helm_fwd_row<NF, 0>(h, pr, j, N, shp, hs, NP, P, s, w6, ys);
      }
    }
  }
  cplx w6[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) w6[f] = s.bot[f][0];
  helm_fwd_row<NF, 1>(h, pr, N - 6, N, shp, hs, NP, P, s, w6, ys);
#pragma unroll
  for (int f = 0; f < NF; ++f) w6[f] = s.bot[f][1];
  helm_fwd_row<NF, 2>(h, pr, N - 5, N, shp, hs, NP, P, s, w6, ys);
#pragma unroll
  for (int f = 0; f < NF; ++f) w6[f] = s.bot[f][2];
  helm_fwd_row<NF, 3>(h, pr, N - 4, N, shp, hs, NP, P, s, w6, ys);
  // border factor rows
  const double DN3 = h.a0 - h.See;
  const double dN3 = 1.0 / DN3;
  const cplx LfN3 = crmul(dN3, csub(mkc(h.a1, 0.0), h.Sfe));
  const cplx LgN3 = crmul(dN3, csub(mkc(h.a2, 0.0), h.Sge));
  h.Sff += cabs2(LfN3) * DN3;
  const double DN2 = h.a0 - h.Sff;
  const double dN2 = 1.0 / DN2;
  h.Sgf = cadd(h.Sgf, cmulc(crmul(DN3, LgN3), LfN3));
  const cplx LgN2 = crmul(dN2, csub(mkc(h.a1, 0.0), h.Sgf));
  h.Sgg += cabs2(LgN3) * DN3;
  h.Sgg += cabs2(LgN2) * DN2;
  const double DN1 = h.a0 - h.Sgg;
  tail.Dm1[0] = dN3;
  tail.Dm1[1] = dN2;
  tail.Dm1[2] = 1.0 / DN1;
  tail.LfN3 = LfN3;
  tail.LgN3 = LgN3;
  tail.LgN2 = LgN2;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    s.w[f][6] = s.top[f][0];
    const cplx yN3 = csub(stencil7<false>(pr, s.w[f]), s.E[f]);
    s.F[f] = cadd(s.F[f], cmul(LfN3, yN3));
    s.G[f] = cadd(s.G[f], cmul(LgN3, yN3));
    window_shift(s.w[f]);
    s.w[f][6] = s.top[f][1];
    const cplx yN2 = csub(stencil7<false>(pr, s.w[f]), s.F[f]);
    s.G[f] = cadd(s.G[f], cmul(LgN2, yN2));
    window_shift(s.w[f]);
    s.w[f][6] = s.top[f][2];
    const cplx yN1 = csub(stencil7<false>(pr, s.w[f]), s.G[f]);
    gst(ys[f] + (long)(N - 3) * P, yN3);
    gst(ys[f] + (long)(N - 2) * P, yN2);
    gst(ys[f] + (long)(N - 1) * P, yN1);
  }
}

struct HRow {   // factor row of the backward Helmholtz sweep
  double dm, La, Lb, Lc;
  cplx le, lf, lg;
};

__device__ __forceinline__ HRow hrow_load(const HScratch& hs, long o, long NP) {
  HRow r;
  r.dm = gldd(hs.hr + o);
  r.La = gldd(hs.hr + NP + o);
  r.Lb = gldd(hs.hr + 2 * NP + o);
  r.Lc = gldd(hs.hr + 3 * NP + o);
  r.le = gld(hs.hc + o);
  r.lf = gld(hs.hc + NP + o);
  r.lg = gld(hs.hc + 2 * NP + o);
  return r;
}

template <int NF, int D, class Sink>
__device__ __forceinline__ void helm_backward(int N, cplx* const* ys, long P, HScratch hs,
                                              const HTail& tail, Sink& sink) {
  BwdState<NF> s;
  const long NP = (long)N * P;
  cplx pyr[NF][D];
  HRow fr[D];
  typename Sink::In pin_[D];
#pragma unroll
  for (int q = 0; q < D; ++q) {
    const int r = N - 4 - q;
#pragma unroll
    for (int f = 0; f < NF; ++f) pyr[f][q] = gld(ys[f] + (long)r * P);
    fr[q] = hrow_load(hs, (long)r * P, NP);
    pin_[q] = sink.load(r);
  }
  {
    const typename Sink::In t1 = sink.load(N - 1), t2 = sink.load(N - 2), t3 = sink.load(N - 3);
    const cplx cg2 = cconj(tail.LgN2), cf3 = cconj(tail.LfN3), cg3 = cconj(tail.LgN3);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const cplx yN1 = gld(ys[f] + (long)(N - 1) * P);
      const cplx yN2 = gld(ys[f] + (long)(N - 2) * P);
      const cplx yN3 = gld(ys[f] + (long)(N - 3) * P);
      s.xN1[f] = crmul(tail.Dm1[2], yN1);
      s.xN2[f] = csub(crmul(tail.Dm1[1], yN2), cmul(cg2, s.xN1[f]));
      s.xN3[f] = csub(csub(crmul(tail.Dm1[0], yN3), cmul(cf3, s.xN2[f])), cmul(cg3, s.xN1[f]));
      s.x1[f] = s.xN3[f];
      s.x2[f] = s.xN2[f];
      s.x3[f] = s.xN1[f];
    }
    sink.apply(N - 1, t1, s.xN1);
    sink.apply(N - 2, t2, s.xN2);
    sink.apply(N - 3, t3, s.xN3);
  }
  for (int ib = N - 4; ib >= 0; ib -= D) {
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const int i = ib - q;
      if (i >= 0) {
        const HRow h = fr[q];
        const typename Sink::In in = pin_[q];
        cplx y[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) y[f] = pyr[f][q];
        const int rn = i - D;
        if (rn >= 0) {
#pragma unroll
          for (int f = 0; f < NF; ++f) pyr[f][q] = gld(ys[f] + (long)rn * P);
          fr[q] = hrow_load(hs, (long)rn * P, NP);
          pin_[q] = sink.load(rn);
        }
        bwd_apply<NF>(i, h.dm, h.La, h.Lb, h.Lc, cconj(h.le), cconj(h.lf), cconj(h.lg), y, in,
                      s, sink);
      }
    }
  }
}

// Sink storing x into NF output fields.
struct NoIn {};

template <int NF>
struct StoreSink {
  typedef NoIn In;
  cplx* const* out;
  long P;
  __device__ In load(int) const { return In(); }
  __device__ void apply(int j, const In&, const cplx* x) {
#pragma unroll
    for (int f = 0; f < NF; ++f) gst(out[f] + (long)j * P, x[f]);
  }
};
