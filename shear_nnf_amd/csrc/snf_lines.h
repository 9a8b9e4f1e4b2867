// Batched y-line solvers: one thread owns one (kx,kz) line of my points.
//
// Storage is [y][kx][kz] complex, so element j of line L = i*mz + k sits at
// j*P + L (P = kxn*mz): consecutive threads read consecutive addresses at
// every j (coalesced), and the serial dependence runs along j inside a
// thread.  The forward substitution writes its intermediate y(j) to a
// scratch field that the backward substitution reads back in reverse.
//
// Algorithms restated from diffy_shear_hepta.f90 (reference):
//   RHS stencil with shear wrap     derivyc 358-383 / lapcdy 487-514
//   L D L^H of the twisted cyclic   modChol7cyclic_cmplx 578-755
//   twisted re-factor of borders    derivyc 389-406 + modChol_addshear 760-870
//   substitution                    solve7cyclic 900-954
#pragma once
#include "snf_internal.h"

// ---------------------------------------------------------------------------
// twisted derivative operator seen by one line
// ---------------------------------------------------------------------------
struct DOp {
  BaseFactor b;
  TwistTail t;
  cplx shp;
  int N;
};

// select t[q] for q in {0,1,2} without dynamic register indexing
__device__ __forceinline__ cplx sel3(const cplx* t, int q) {
  return q == 0 ? t[0] : (q == 1 ? t[1] : t[2]);
}
__device__ __forceinline__ cplx dop_Le(const DOp& o, int j) {
  return (j < o.N - 6) ? crmul(o.b.Ler[j], o.shp) : sel3(o.t.Le, j - (o.N - 6));
}
__device__ __forceinline__ cplx dop_Lf(const DOp& o, int j) {
  return (j < o.N - 5) ? crmul(o.b.Lfr[j], o.shp) : sel3(o.t.Lf, j - (o.N - 5));
}
__device__ __forceinline__ cplx dop_Lg(const DOp& o, int j) {
  return (j < o.N - 4) ? crmul(o.b.Lgr[j], o.shp) : sel3(o.t.Lg, j - (o.N - 4));
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
// top[0..2] = u(N..N+2) = u(0..2)*shp.
__device__ __forceinline__ void window_init(const cplx* __restrict__ in, long P, int N,
                                            cplx shp, cplx shm, cplx* w, cplx* top) {
  w[0] = cmul(in[(long)(N - 3) * P], shm);
  w[1] = cmul(in[(long)(N - 2) * P], shm);
  w[2] = cmul(in[(long)(N - 1) * P], shm);
  w[3] = in[0];
  w[4] = in[P];
  w[5] = in[2 * P];
  top[0] = cmul(w[3], shp);
  top[1] = cmul(w[4], shp);
  top[2] = cmul(w[5], shp);
}

__device__ __forceinline__ void window_shift(cplx* w) {
#pragma unroll
  for (int m = 0; m < 6; ++m) w[m] = w[m + 1];
}

// ---------------------------------------------------------------------------
// derivative solve (twisted derivative factor shared by NF fields)
//   forward:  y = L^-1 (R u)   -> ys
//   backward: x = L^-H D^-1 y  -> sink(j, x[])
// ---------------------------------------------------------------------------
template <int NF, bool FIRST>
__device__ __forceinline__ void deriv_forward(const DOp& o, const double* c,
                                              const cplx* const* in, long pin,
                                              cplx* const* ys, long py, cplx shm) {
  const int N = o.N;
  cplx w[NF][7], top[NF][3];
  cplx y1[NF], y2[NF], y3[NF], E[NF], F[NF], G[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    window_init(in[f], pin, N, o.shp, shm, w[f], top[f]);
    y1[f] = y2[f] = y3[f] = E[f] = F[f] = G[f] = mkc(0.0, 0.0);
  }
  // interior rows 0..N-4: u(j+3) is always inside the line
#pragma unroll 2
  for (int j = 0; j <= N - 4; ++j) {
    const double La = o.b.La[j], Lb = o.b.Lb[j], Lc = o.b.Lc[j];
    const cplx le = dop_Le(o, j), lf = dop_Lf(o, j), lg = dop_Lg(o, j);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      w[f][6] = in[f][(long)(j + 3) * pin];
      const cplx b = stencil7<FIRST>(c, w[f]);
      cplx y;
      y.x = b.x - (Lc * y3[f].x + Lb * y2[f].x + La * y1[f].x);
      y.y = b.y - (Lc * y3[f].y + Lb * y2[f].y + La * y1[f].y);
      E[f] = cadd(E[f], cmul(le, y));
      F[f] = cadd(F[f], cmul(lf, y));
      G[f] = cadd(G[f], cmul(lg, y));
      ys[f][(long)j * py] = y;
      y3[f] = y2[f];
      y2[f] = y1[f];
      y1[f] = y;
      window_shift(w[f]);
    }
  }
  // border rows N-3, N-2, N-1 (wrap with shp at the top)
  const cplx lfN3 = dop_Lf(o, N - 3), lgN3 = dop_Lg(o, N - 3), lgN2 = dop_Lg(o, N - 2);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    w[f][6] = top[f][0];
    cplx yN3 = csub(stencil7<FIRST>(c, w[f]), E[f]);
    F[f] = cadd(F[f], cmul(lfN3, yN3));
    G[f] = cadd(G[f], cmul(lgN3, yN3));
    window_shift(w[f]);
    w[f][6] = top[f][1];
    cplx yN2 = csub(stencil7<FIRST>(c, w[f]), F[f]);
    G[f] = cadd(G[f], cmul(lgN2, yN2));
    window_shift(w[f]);
    w[f][6] = top[f][2];
    cplx yN1 = csub(stencil7<FIRST>(c, w[f]), G[f]);
    ys[f][(long)(N - 3) * py] = yN3;
    ys[f][(long)(N - 2) * py] = yN2;
    ys[f][(long)(N - 1) * py] = yN1;
  }
}

template <int NF, class Sink>
__device__ __forceinline__ void deriv_backward(const DOp& o, cplx* const* ys, long P,
                                               Sink& sink) {  // P: stride of ys
  const int N = o.N;
  cplx xN1[NF], xN2[NF], xN3[NF], x1[NF], x2[NF], x3[NF];
  const cplx cg2 = cconj(dop_Lg(o, N - 2));
  const cplx cf3 = cconj(dop_Lf(o, N - 3)), cg3 = cconj(dop_Lg(o, N - 3));
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const cplx yN1 = ys[f][(long)(N - 1) * P];
    const cplx yN2 = ys[f][(long)(N - 2) * P];
    const cplx yN3 = ys[f][(long)(N - 3) * P];
    xN1[f] = crmul(o.t.Dm1[2], yN1);
    xN2[f] = csub(crmul(o.t.Dm1[1], yN2), cmul(cg2, xN1[f]));
    xN3[f] = csub(csub(crmul(o.t.Dm1[0], yN3), cmul(cf3, xN2[f])), cmul(cg3, xN1[f]));
    x1[f] = xN3[f];
    x2[f] = xN2[f];
    x3[f] = xN1[f];
  }
  sink(N - 1, xN1);
  sink(N - 2, xN2);
  sink(N - 3, xN3);
#pragma unroll 2
  for (int i = N - 4; i >= 0; --i) {
    const double dm = o.b.Dm1[i];
    const double La = o.b.La[i + 1], Lb = o.b.Lb[i + 2], Lc = o.b.Lc[i + 3];
    const cplx ce = cconj(dop_Le(o, i)), cf = cconj(dop_Lf(o, i)), cg = cconj(dop_Lg(o, i));
    cplx xo[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const cplx y = ys[f][(long)i * P];
      cplx x = crmul(dm, y);
      x.x = x.x - La * x1[f].x - Lb * x2[f].x - Lc * x3[f].x;
      x.y = x.y - La * x1[f].y - Lb * x2[f].y - Lc * x3[f].y;
      x = csub(csub(csub(x, cmul(ce, xN3[f])), cmul(cf, xN2[f])), cmul(cg, xN1[f]));
      x3[f] = x2[f];
      x2[f] = x1[f];
      x1[f] = x;
      xo[f] = x;
    }
    sink(i, xo);
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

template <int NF>
__device__ __forceinline__ void helm_forward(const double* a /*7: lap matrix*/, const double* pr,
                                             int N, cplx shp, cplx shm,
                                             const cplx* const* in, long pin,
                                             cplx* const* ys, long P, HScratch hs,
                                             HTail& tail) {  // P: stride of ys and hs
  const double a0 = a[3], a1 = a[4], a2 = a[5], a3 = a[6];
  cplx w[NF][7], top[NF][3];
  cplx y1[NF], y2[NF], y3[NF], E[NF], F[NF], G[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    window_init(in[f], pin, N, shp, shm, w[f], top[f]);
    y1[f] = y2[f] = y3[f] = E[f] = F[f] = G[f] = mkc(0.0, 0.0);
  }
  double D1 = 0.0, D2 = 0.0, D3 = 0.0;   // D(j-1), D(j-2), D(j-3)
  double La_j = 0.0, Lb_j = 0.0, Lb_j1 = 0.0, Lc_j = 0.0, Lc_j1 = 0.0, Lc_j2 = 0.0;
  cplx e1 = mkc(0, 0), e2 = e1, e3 = e1, f1 = e1, f2 = e1, f3 = e1, g1 = e1, g2 = e1, g3 = e1;
  double See = 0.0, Sff = 0.0, Sgg = 0.0;
  cplx Sfe = mkc(0, 0), Sge = Sfe, Sgf = Sfe;
  const int last = N - 4;
  for (int j = 0; j <= last; ++j) {
    // interior factor row j
    const double D = a0 - (Lc_j * D3 * Lc_j + Lb_j * D2 * Lb_j + La_j * D1 * La_j);
    const double dm = 1.0 / D;
    const double La_n = (j + 1 <= last) ? dm * (a1 - (Lc_j1 * D2 * Lb_j + Lb_j1 * D1 * La_j)) : 0.0;
    const double Lb_n = (j + 2 <= last) ? dm * (a2 - Lc_j2 * D1 * La_j) : 0.0;
    const double Lc_n = (j + 3 <= last) ? dm * a3 : 0.0;
    // border rows (entries A(row, j) of the cyclic matrix)
    cplx se = mkc(0, 0), sf = se, sg = se;
    if (j == 0) { se = crmul(a3, shp); sf = crmul(a2, shp); sg = crmul(a1, shp); }
    if (j == 1) { sf = crmul(a3, shp); sg = crmul(a2, shp); }
    if (j == 2) { sg = crmul(a3, shp); }
    if (j == N - 6) se.x += a3;
    if (j == N - 5) { se.x += a2; sf.x += a3; }
    if (j == N - 4) { se.x += a1; sf.x += a2; sg.x += a3; }
    cplx le, lf, lg;
    le.x = dm * (se.x - e3.x * D3 * Lc_j - e2.x * D2 * Lb_j - e1.x * D1 * La_j);
    le.y = dm * (se.y - e3.y * D3 * Lc_j - e2.y * D2 * Lb_j - e1.y * D1 * La_j);
    lf.x = dm * (sf.x - f3.x * D3 * Lc_j - f2.x * D2 * Lb_j - f1.x * D1 * La_j);
    lf.y = dm * (sf.y - f3.y * D3 * Lc_j - f2.y * D2 * Lb_j - f1.y * D1 * La_j);
    lg.x = dm * (sg.x - g3.x * D3 * Lc_j - g2.x * D2 * Lb_j - g1.x * D1 * La_j);
    lg.y = dm * (sg.y - g3.y * D3 * Lc_j - g2.y * D2 * Lb_j - g1.y * D1 * La_j);
    See += cabs2(le) * D;
    Sfe = cadd(Sfe, crmul(D, cmulc(lf, le)));
    Sff += cabs2(lf) * D;
    Sge = cadd(Sge, cmulc(crmul(D, lg), le));
    Sgf = cadd(Sgf, cmulc(crmul(D, lg), lf));
    Sgg += cabs2(lg) * D;
    // forward substitution of row j
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      w[f][6] = in[f][(long)(j + 3) * pin];
      const cplx b = stencil7<false>(pr, w[f]);
      cplx y;
      y.x = b.x - (Lc_j * y3[f].x + Lb_j * y2[f].x + La_j * y1[f].x);
      y.y = b.y - (Lc_j * y3[f].y + Lb_j * y2[f].y + La_j * y1[f].y);
      E[f] = cadd(E[f], cmul(le, y));
      F[f] = cadd(F[f], cmul(lf, y));
      G[f] = cadd(G[f], cmul(lg, y));
      ys[f][(long)j * P] = y;
      y3[f] = y2[f];
      y2[f] = y1[f];
      y1[f] = y;
      window_shift(w[f]);
    }
    const long o = (long)j * P;
    const long NP = (long)N * P;
    hs.hr[o] = dm;
    hs.hr[NP + o] = La_n;
    hs.hr[2 * NP + o] = Lb_n;
    hs.hr[3 * NP + o] = Lc_n;
    hs.hc[o] = le;
    hs.hc[NP + o] = lf;
    hs.hc[2 * NP + o] = lg;
    // advance the recurrences
    D3 = D2; D2 = D1; D1 = D;
    La_j = La_n;
    Lb_j = Lb_j1; Lb_j1 = Lb_n;
    Lc_j = Lc_j1; Lc_j1 = Lc_j2; Lc_j2 = Lc_n;
    e3 = e2; e2 = e1; e1 = le;
    f3 = f2; f2 = f1; f1 = lf;
    g3 = g2; g2 = g1; g1 = lg;
  }
  // border factor rows
  const double DN3 = a0 - See;
  const double dN3 = 1.0 / DN3;
  cplx LfN3 = crmul(dN3, csub(mkc(a1, 0.0), Sfe));
  cplx LgN3 = crmul(dN3, csub(mkc(a2, 0.0), Sge));
  Sff += cabs2(LfN3) * DN3;
  const double DN2 = a0 - Sff;
  const double dN2 = 1.0 / DN2;
  Sgf = cadd(Sgf, cmulc(crmul(DN3, LgN3), LfN3));
  cplx LgN2 = crmul(dN2, csub(mkc(a1, 0.0), Sgf));
  Sgg += cabs2(LgN3) * DN3;
  Sgg += cabs2(LgN2) * DN2;
  const double DN1 = a0 - Sgg;
  tail.Dm1[0] = dN3;
  tail.Dm1[1] = dN2;
  tail.Dm1[2] = 1.0 / DN1;
  tail.LfN3 = LfN3;
  tail.LgN3 = LgN3;
  tail.LgN2 = LgN2;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    w[f][6] = top[f][0];
    cplx yN3 = csub(stencil7<false>(pr, w[f]), E[f]);
    F[f] = cadd(F[f], cmul(LfN3, yN3));
    G[f] = cadd(G[f], cmul(LgN3, yN3));
    window_shift(w[f]);
    w[f][6] = top[f][1];
    cplx yN2 = csub(stencil7<false>(pr, w[f]), F[f]);
    G[f] = cadd(G[f], cmul(LgN2, yN2));
    window_shift(w[f]);
    w[f][6] = top[f][2];
    cplx yN1 = csub(stencil7<false>(pr, w[f]), G[f]);
    ys[f][(long)(N - 3) * P] = yN3;
    ys[f][(long)(N - 2) * P] = yN2;
    ys[f][(long)(N - 1) * P] = yN1;
  }
}

template <int NF, class Sink>
__device__ __forceinline__ void helm_backward(int N, cplx* const* ys, long P, HScratch hs,
                                              const HTail& tail, Sink& sink) {
  cplx xN1[NF], xN2[NF], xN3[NF], x1[NF], x2[NF], x3[NF];
  const cplx cg2 = cconj(tail.LgN2), cf3 = cconj(tail.LfN3), cg3 = cconj(tail.LgN3);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const cplx yN1 = ys[f][(long)(N - 1) * P];
    const cplx yN2 = ys[f][(long)(N - 2) * P];
    const cplx yN3 = ys[f][(long)(N - 3) * P];
    xN1[f] = crmul(tail.Dm1[2], yN1);
    xN2[f] = csub(crmul(tail.Dm1[1], yN2), cmul(cg2, xN1[f]));
    xN3[f] = csub(csub(crmul(tail.Dm1[0], yN3), cmul(cf3, xN2[f])), cmul(cg3, xN1[f]));
    x1[f] = xN3[f];
    x2[f] = xN2[f];
    x3[f] = xN1[f];
  }
  sink(N - 1, xN1);
  sink(N - 2, xN2);
  sink(N - 3, xN3);
  const long NP = (long)N * P;
#pragma unroll 2
  for (int i = N - 4; i >= 0; --i) {
    const long o = (long)i * P;
    const double dm = hs.hr[o];
    const double La = hs.hr[NP + o], Lb = hs.hr[2 * NP + o], Lc = hs.hr[3 * NP + o];
    const cplx ce = cconj(hs.hc[o]), cf = cconj(hs.hc[NP + o]), cg = cconj(hs.hc[2 * NP + o]);
    cplx xo[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const cplx y = ys[f][o];
      cplx x = crmul(dm, y);
      x.x = x.x - La * x1[f].x - Lb * x2[f].x - Lc * x3[f].x;
      x.y = x.y - La * x1[f].y - Lb * x2[f].y - Lc * x3[f].y;
      x = csub(csub(csub(x, cmul(ce, xN3[f])), cmul(cf, xN2[f])), cmul(cg, xN1[f]));
      x3[f] = x2[f];
      x2[f] = x1[f];
      x1[f] = x;
      xo[f] = x;
    }
    sink(i, xo);
  }
}

// Simple sink: store x into an output field.
template <int NF>
struct StoreSink {
  cplx* const* out;
  long P;
  __device__ void operator()(int j, const cplx* x) {
#pragma unroll
    for (int f = 0; f < NF; ++f) out[f][(long)j * P] = x[f];
  }
};
