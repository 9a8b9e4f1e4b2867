// y-line stages of one RK3 substep (cross.f90:428-1000).
//
// Explicit substeps use two threads per (kx,kz) line, which the reference
// equations allow because the omega_y and phi updates are independent:
//   vor chain:  d2(vor) -> viscous + RK update of vor, vorwk -> d(vor)/dy
//   phi chain:  d(rdome)/dy and d2(phi) -> RK update of phi, phiwk
//               -> v = lapcdy(phi) -> dv/dy
// The x-z stage outputs (rdome, rhg, rhv) and the y-stage outputs consumed by
// the next x-z stage (v, dome, dv) live in separate fields, so the chains
// never read a line the other chain writes.  Two extra threads advance the
// real x-z means U(y) = u00 and W(y) = w00 with the untwisted operators
// (m=1 calls of derivyc / visc3d / lapcdy).
#include "snf_internal.h"
#include "snf_lines.h"
#include "snf_yfft.h"
#include "snf_yargs.h"
#include "snf_yfk.h"

// ---------------------------------------------------------------------------
// two twisted derivative operators of the same line advanced in lock step
// (d/dy of one field and d2/dy2 of another, as in cross.f90:565-664); the
// sweeps are independent, and interleaving them keeps both streams in flight
// ---------------------------------------------------------------------------
template <bool FA, bool FB, int D>
__device__ __forceinline__ void deriv_forward_dual(const DOp& oa, const double* ca, const DOp& ob,
                                                   const double* cb, const cplx* ina,
                                                   const cplx* inb, long pin, cplx* ysa,
                                                   cplx* ysb, long py, cplx shm) {
  const int N = oa.N;
  const int mlast = N - 7;
  FwdState<1, FA> sa;
  FwdState<1, FB> sb;
  window_init(ina, pin, N, oa.shp, shm, sa.w[0], sa.bot[0], sa.top[0]);
  window_init(inb, pin, N, ob.shp, shm, sb.w[0], sb.bot[0], sb.top[0]);
  sa.y1[0] = sa.y2[0] = sa.y3[0] = sa.E[0] = sa.F[0] = sa.G[0] = mkc(0, 0);
  sb.y1[0] = sb.y2[0] = sb.y3[0] = sb.E[0] = sb.F[0] = sb.G[0] = mkc(0, 0);
  cplx pa[D], pb[D];
  cplx* ya[1] = {ysa};
  cplx* yb[1] = {ysb};
#pragma unroll
  for (int q = 0; q < D; ++q) {
    pa[q] = gld(ina + (long)(3 + q) * pin);
    pb[q] = gld(inb + (long)(3 + q) * pin);
  }
  for (int jb = 0; jb <= mlast; jb += D) {
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const int j = jb + q;
      if (j <= mlast) {
        const int rn = j + 3 + D;
        const cplx wa = pa[q], wb = pb[q];
        if (rn <= N - 4) {
          pa[q] = gld(ina + (long)rn * pin);
          pb[q] = gld(inb + (long)rn * pin);
        }
        deriv_fwd_row<1, FA, 0>(oa, ca, j, sa, &wa, ya, py);
        deriv_fwd_row<1, FB, 0>(ob, cb, j, sb, &wb, yb, py);
      }
    }
  }
  deriv_fwd_row<1, FA, 1>(oa, ca, N - 6, sa, &sa.bot[0][0], ya, py);
  deriv_fwd_row<1, FB, 1>(ob, cb, N - 6, sb, &sb.bot[0][0], yb, py);
  deriv_fwd_row<1, FA, 2>(oa, ca, N - 5, sa, &sa.bot[0][1], ya, py);
  deriv_fwd_row<1, FB, 2>(ob, cb, N - 5, sb, &sb.bot[0][1], yb, py);
  deriv_fwd_row<1, FA, 3>(oa, ca, N - 4, sa, &sa.bot[0][2], ya, py);
  deriv_fwd_row<1, FB, 3>(ob, cb, N - 4, sb, &sb.bot[0][2], yb, py);
  border_rows<1, FA>(oa, ca, sa, ya, py);
  border_rows<1, FB>(ob, cb, sb, yb, py);
}

// Backward sweeps of two operators; the sink receives x = [xa, xb].
template <int D, class Sink>
__device__ __forceinline__ void deriv_backward_dual(const DOp& oa, const DOp& ob, const cplx* ysa,
                                                    const cplx* ysb, long py, Sink& sink) {
  const int N = oa.N;
  const int mtop = N - 7;
  BwdState<1> sa, sb;
  auto border = [&](const DOp& o, const cplx* ys, BwdState<1>& s) {
    const cplx cg2 = cconj(o.lg2), cf3 = cconj(o.lf2), cg3 = cconj(o.lg1);
    const cplx yN1 = gld(ys + (long)(N - 1) * py);
    const cplx yN2 = gld(ys + (long)(N - 2) * py);
    const cplx yN3 = gld(ys + (long)(N - 3) * py);
    s.xN1[0] = crmul(o.dm2, yN1);
    s.xN2[0] = csub(crmul(o.dm1, yN2), cmul(cg2, s.xN1[0]));
    s.xN3[0] = csub(csub(crmul(o.dm0, yN3), cmul(cf3, s.xN2[0])), cmul(cg3, s.xN1[0]));
    s.x1[0] = s.xN3[0];
    s.x2[0] = s.xN2[0];
    s.x3[0] = s.xN1[0];
  };
  cplx pa[D], pb[D];
  typename Sink::In pin_[D];
#pragma unroll
  for (int q = 0; q < D; ++q) {
    const int r = mtop - q;
    pa[q] = gld(ysa + (long)r * py);
    pb[q] = gld(ysb + (long)r * py);
    pin_[q] = sink.load(r);
  }
  {
    const typename Sink::In t1 = sink.load(N - 1), t2 = sink.load(N - 2), t3 = sink.load(N - 3);
    border(oa, ysa, sa);
    border(ob, ysb, sb);
    cplx x2[2];
    x2[0] = sa.xN1[0]; x2[1] = sb.xN1[0];
    sink.apply(N - 1, t1, x2);
    x2[0] = sa.xN2[0]; x2[1] = sb.xN2[0];
    sink.apply(N - 2, t2, x2);
    x2[0] = sa.xN3[0]; x2[1] = sb.xN3[0];
    sink.apply(N - 3, t3, x2);
  }
  // one row of both sweeps with explicit border coefficients
  auto row = [&](int i, const cplx& ya, const cplx& yb, const typename Sink::In& in,
                 cplx cea, cplx cfa, cplx cga, cplx ceb, cplx cfb, cplx cgb) {
    cplx xo[2];
    xo[0] = bwd_row(ya, cld(oa.b.Dm1 + i), cld(oa.b.La + i + 1), cld(oa.b.Lb + i + 2),
                    cld(oa.b.Lc + i + 3), cea, cfa, cga, sa.xN3[0], sa.xN2[0], sa.xN1[0],
                    sa.x1[0], sa.x2[0], sa.x3[0]);
    xo[1] = bwd_row(yb, cld(ob.b.Dm1 + i), cld(ob.b.La + i + 1), cld(ob.b.Lb + i + 2),
                    cld(ob.b.Lc + i + 3), ceb, cfb, cgb, sb.xN3[0], sb.xN2[0], sb.xN1[0],
                    sb.x1[0], sb.x2[0], sb.x3[0]);
    sink.apply(i, in, xo);
  };
  {
    int i = N - 4;
    row(i, gld(ysa + (long)i * py), gld(ysb + (long)i * py), sink.load(i), cconj(oa.le2),
        cconj(oa.lf1), cconj(oa.lg0), cconj(ob.le2), cconj(ob.lf1), cconj(ob.lg0));
    i = N - 5;
    row(i, gld(ysa + (long)i * py), gld(ysb + (long)i * py), sink.load(i), cconj(oa.le1),
        cconj(oa.lf0), cconj(dop_g(oa, i)), cconj(ob.le1), cconj(ob.lf0), cconj(dop_g(ob, i)));
    i = N - 6;
    row(i, gld(ysa + (long)i * py), gld(ysb + (long)i * py), sink.load(i), cconj(oa.le0),
        cconj(dop_f(oa, i)), cconj(dop_g(oa, i)), cconj(ob.le0), cconj(dop_f(ob, i)),
        cconj(dop_g(ob, i)));
  }
  for (int ib = mtop; ib >= 0; ib -= D) {
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const int i = ib - q;
      if (i >= 0) {
        const typename Sink::In in = pin_[q];
        const cplx ya = pa[q], yb = pb[q];
        const int rn = i - D;
        if (rn >= 0) {
          pa[q] = gld(ysa + (long)rn * py);
          pb[q] = gld(ysb + (long)rn * py);
          pin_[q] = sink.load(rn);
        }
        row(i, ya, yb, in, cconj(dop_e(oa, i)), cconj(dop_f(oa, i)), cconj(dop_g(oa, i)),
            cconj(dop_e(ob, i)), cconj(dop_f(ob, i)), cconj(dop_g(ob, i)));
      }
    }
  }
}

// lap matrix rh2 - k2 pr2 (lapcdy, diffy_shear_hepta.f90:475)
__device__ __forceinline__ void lapmat(const YCoef& c, double k2, double* a) {
#pragma unroll
  for (int m = 0; m < 7; ++m) a[m] = c.rh2[m] - k2 * c.pr2[m];
}

__device__ __forceinline__ void line_geom(const Geom& g, long L, int& i, int& k, double& kx,
                                          double& k2) {
  i = (int)(L / g.mz);
  k = (int)(L - (long)i * g.mz);
  kx = g.alp * (double)i;
  const double kz = kz_of(k, g.nz, g.mz, g.gam);
  k2 = kx * kx + kz * kz;
}

template <int D>
__device__ __forceinline__ void d1_one(const YArgs& a, const DOp& o, cplx shm, const cplx* in,
                                       long pin, cplx* ys, cplx* out, long pout) {
  const cplx* ins[1] = {in};
  cplx* yy[1] = {ys};
  deriv_forward<1, true, D>(o, a.c.rh1, ins, pin, yy, a.PS, shm);
  cplx* outs[1] = {out};
  StoreSink<1> sk{outs, pout};
  deriv_backward<1, D>(o, yy, a.PS, sk);
}

__device__ __forceinline__ void helm_line(const YArgs& a, long L, double k2, cplx shp, cplx shm,
                                          const cplx* in, long pin, cplx* ys, cplx* out,
                                          long pout) {
  double lm[7];
  lapmat(a.c, k2, lm);
  HScratch hs{a.hs.hr + L, a.hs.hc + L};
  HTail tail;
  const cplx* ins[1] = {in};
  cplx* yy[1] = {ys};
  helm_forward<1, 8>(lm, a.c.pr2, a.g.my, shp, shm, ins, pin, yy, a.PS, hs, tail);
  cplx* outs[1] = {out};
  StoreSink<1> sk{outs, pout};
  helm_backward<1, 4>(a.g.my, yy, a.PS, hs, tail, sk);
}

struct RealOutSink {  // dd[j] = Re x
  typedef NoIn In;
  double* dd;
  __device__ In load(int) const { return In(); }
  __device__ void apply(int j, const In&, const cplx* x) { gstd(dd + j, x[0].x); }
};

// D1 of a 00 line (real periodic, derivyc m=1) -> dd
__device__ __forceinline__ void d1_real(const YArgs& a, long L, const cplx* u, double* dd) {
  const DOp o = make_dop(a.b1, a.id1, mkc(1.0, 0.0), a.g.my);
  const cplx* in[1] = {u};
  cplx* ys[1] = {a.s1 + L};
  deriv_forward<1, true, 4>(o, a.c.rh1, in, 1, ys, a.PS, mkc(1.0, 0.0));
  RealOutSink sk{dd};
  deriv_backward<1, 4>(o, ys, a.PS, sk);
}

// ---------------------------------------------------------------------------
// first step of an advance call: v = lapcdy(phi), then D1(vor), D1(v)
// (cross.f90:408-426 and 434-445); 00 lines: dU/dy, dW/dy (cross.f90:1198-1205)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(128)
k_y_first(YArgs a) {
  const long L = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long P = a.g.P;
  if (L >= P + 2) return;
  if (a.sc->halted) return;
  if (L < P) {
    int i, k;
    double kx, k2;
    line_geom(a.g, L, i, k, kx, k2);
    const long LB = sline(a.g, i, k), FS = SNF_CW;   // the line in the fields
    if (L != 0) helm_line(a, L, k2, gld(a.shp_old + i), gld(a.shm_old + i), a.phi + LB, FS, a.s2 + L, a.v + LB, FS);
    const DOp o = make_dop(a.b1, a.tw_old + i, gld(a.shp_old + i), a.g.my);
    d1_one<8>(a, o, gld(a.shm_old + i), a.vor + LB, FS, a.s1 + L, a.dome + LB, FS);
    d1_one<8>(a, o, gld(a.shm_old + i), a.v + LB, FS, a.s2 + L, a.dv + LB, FS);
  } else if (L == P) {
    d1_real(a, L, a.u00c, a.dudy);
  } else {
    d1_real(a, L, a.w00c, a.dwdy);
  }
}

// dU/dy and dW/dy only (top of hvhg, cross.f90:1198-1205)
__global__ void __launch_bounds__(64)
k_d1_00(YArgs a) {
  if (threadIdx.x == 0) d1_real(a, a.g.P, a.u00c, a.dudy);
  if (threadIdx.x == 1) d1_real(a, a.g.P + 1, a.w00c, a.dwdy);
}

// ---------------------------------------------------------------------------
// explicit substep (cross.f90:565-796)
// ---------------------------------------------------------------------------

struct C3 {
  cplx a, b, c;
};

// vor chain: hg + nu (d2 vor - k2 vor) (visc3d), RK accumulation, update,
// zeta*hg with the phase map, un-mapping of vor (cross.f90:652-781)
struct VorSink {
  typedef C3 In;   // rhg, vor, vorwk
  const cplx *rhg;
  cplx *vor, *vorwk;
  long P;
  double fac, k2, kx, dt, gam, zet, f1, f2, dy, Ly;
  int rk, is00;
  __device__ In load(int j) const {
    const long o = (long)j * P;
    return In{gld(rhg + o), gld(vor + o), rk == 1 ? mkc(0, 0) : gld(vorwk + o)};
  }
  __device__ void apply(int j, const In& in, const cplx* x) {
    const long o = (long)j * P;
    cplx hgj = in.a;
    if (!is00) hgj = cadd(hgj, crmul(fac, cadd(crmul(-k2, in.b), x[0])));
    cplx vwk = (rk == 1) ? crmul(gam, hgj) : cadd(in.c, crmul(gam, hgj));
    const cplx vn = cadd(in.b, crmul(dt, vwk));
    const double yj = (double)j * dy - 0.5 * Ly;
    if (rk != 3) vwk = cmul(crmul(zet, hgj), mframe(kx, yj, f1));
    gst(vor + o, cmul(vn, mframe(kx, yj, f2)));
    gst(vorwk + o, vwk);
  }
};

// phi chain: hv - d(rdome)/dy, + nu (d2 phi - k2 phi), RK update, maps
// (cross.f90:607, 652-781)
struct PhiSink {
  typedef C3 In;   // rhv, phi, phiwk
  const cplx* rhv;
  cplx *phi, *phiwk;
  long P;
  double fac, k2, kx, dt, gam, zet, f1, f2, dy, Ly;
  int rk, is00;
  __device__ In load(int j) const {
    const long o = (long)j * P;
    return In{gld(rhv + o), gld(phi + o), rk == 1 ? mkc(0, 0) : gld(phiwk + o)};
  }
  __device__ void apply(int j, const In& in, const cplx* x) {
    const long o = (long)j * P;
    cplx hvj = csub(in.a, x[0]);
    if (!is00) hvj = cadd(hvj, crmul(fac, cadd(crmul(-k2, in.b), x[1])));
    cplx pwk = (rk == 1) ? crmul(gam, hvj) : cadd(in.c, crmul(gam, hvj));
    const cplx pn = cadd(in.b, crmul(dt, pwk));
    const double yj = (double)j * dy - 0.5 * Ly;
    if (rk != 3) pwk = cmul(crmul(zet, hvj), mframe(kx, yj, f1));
    gst(phi + o, cmul(pn, mframe(kx, yj, f2)));
    gst(phiwk + o, pwk);
  }
};

struct R2 {
  double u, wk;
};

// 00 lines: rf + nu d2u (visc3d m=1, rk=0), RK update of U or W
struct Explicit00Sink {
  typedef R2 In;
  cplx *u, *wk;
  const double* rf;
  double fac, dt, gam, zet;
  int rk;
  __device__ In load(int j) const { return In{gld(u + j).x, rk == 1 ? 0.0 : gld(wk + j).x}; }
  __device__ void apply(int j, const In& in, const cplx* x) {
    const double rfj = gldd(rf + j) + fac * (-0.0 * in.u + x[0].x);
    double wkj = (rk == 1) ? gam * rfj : in.wk + gam * rfj;
    const double un = in.u + dt * wkj;
    if (rk != 3) wkj = rfj * zet;
    gst(u + j, mkc(un, 0.0));
    gst(wk + j, mkc(wkj, 0.0));
  }
};

__global__ void __launch_bounds__(128)
k_y_explicit(YArgs a, int rk) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long P = a.g.P;
  if (t >= 2 * P + 2) return;
  if (a.sc->halted) return;
  const int ir = rk - 1;
  double dt, f1, f2;
  substep_factors(a.sc, a.g.s, rk, dt, f1, f2);
  const double fac = 1.0 / a.g.Re;
  const int N = a.g.my;
  if (t < 2 * P) {
    const bool vchain = t < P;
    const long L = vchain ? t : t - P;
    int i, k;
    double kx, k2;
    line_geom(a.g, L, i, k, kx, k2);
    const long LB = sline(a.g, i, k), FS = SNF_CW;   // the line in the fields
    const DOp o1 = make_dop(a.b1, a.tw_old + i, gld(a.shp_old + i), N);
    const DOp o2 = make_dop(a.b2, a.tw_old + a.g.kxn + i, gld(a.shp_old + i), N);
    const DOp o1n = make_dop(a.b1, a.tw_new + i, gld(a.shp_new + i), N);
    if (vchain) {
      const cplx* in[1] = {a.vor + LB};
      cplx* ys[1] = {a.s1 + L};
      deriv_forward<1, false, 8>(o2, a.c.rh2, in, FS, ys, a.PS, gld(a.shm_old + i));
      VorSink sk{a.rhg + LB, a.vor + LB, a.vorwk + LB, FS, fac, k2, kx, dt, c_gama[ir], c_zeta[ir],
                 f1, f2, a.g.dy, a.g.Ly, rk, L == 0};
      deriv_backward<1, 4>(o2, ys, a.PS, sk);
      d1_one<8>(a, o1n, gld(a.shm_new + i), a.vor + LB, FS, a.s1 + L, a.dome + LB, FS);
    } else {
      deriv_forward_dual<true, false, 8>(o1, a.c.rh1, o2, a.c.rh2, a.rdome + LB, a.phi + LB, FS,
                                         a.s2 + L, a.s3 + L, a.PS, gld(a.shm_old + i));
      PhiSink sk{a.rhv + LB, a.phi + LB, a.phiwk + LB, FS, fac, k2, kx, dt, c_gama[ir], c_zeta[ir],
                 f1, f2, a.g.dy, a.g.Ly, rk, L == 0};
      deriv_backward_dual<4>(o1, o2, a.s2 + L, a.s3 + L, a.PS, sk);
      if (L != 0)
        helm_line(a, L, k2, gld(a.shp_new + i), gld(a.shm_new + i), a.phi + LB, FS, a.s2 + L, a.v + LB, FS);
      d1_one<8>(a, o1n, gld(a.shm_new + i), a.v + LB, FS, a.s2 + L, a.dv + LB, FS);
    }
  } else {
    const long L = P + (t - 2 * P);
    cplx* u = (L == P) ? a.u00c : a.w00c;
    cplx* wk = (L == P) ? a.u00wk : a.w00wk;
    const double* rf = (L == P) ? a.rf0u : a.rf0w;
    double* dd = (L == P) ? a.dudy : a.dwdy;
    const DOp o = make_dop(a.b2, a.id2, mkc(1.0, 0.0), N);
    const cplx* in[1] = {u};
    cplx* ys[1] = {a.s1 + L};
    deriv_forward<1, false, 4>(o, a.c.rh2, in, 1, ys, a.PS, mkc(1.0, 0.0));
    Explicit00Sink sk{u, wk, rf, fac, dt, c_gama[ir], c_zeta[ir], rk};
    deriv_backward<1, 4>(o, ys, a.PS, sk);
    d1_real(a, L, u, dd);
  }
}

// ---------------------------------------------------------------------------
// implicit substep (cross.f90:797-990), one thread per line
// ---------------------------------------------------------------------------
struct SubSink {   // rhv = rhv - x  (hv = hv - dvordy)
  typedef cplx In;
  cplx* out;
  long P;
  __device__ In load(int j) const { return gld(out + (long)j * P); }
  __device__ void apply(int j, const In& in, const cplx* x) { gst(out + (long)j * P, csub(in, x[0])); }
};

struct C6 {
  cplx hg, hv, vor, phi, vwk, pwk;
};

struct ImplicitRKSink {
  typedef C6 In;
  const cplx *rhg, *rhv;
  cplx *vor, *phi, *vorwk, *phiwk;
  long P;
  double fac, facb, rdt, k2, kx, gam, f2, dy, Ly;
  int rk, is00;
  __device__ In load(int j) const {
    const long o = (long)j * P;
    const cplx z = mkc(0, 0);
    return In{gld(rhg + o), gld(rhv + o), gld(vor + o), gld(phi + o), rk == 1 ? z : gld(vorwk + o),
              rk == 1 ? z : gld(phiwk + o)};
  }
  __device__ void apply(int j, const In& in, const cplx* x) {
    const long o = (long)j * P;
    cplx vwk = (rk == 1) ? crmul(gam, in.hg) : cadd(in.vwk, crmul(gam, in.hg));
    cplx pwk = (rk == 1) ? crmul(gam, in.hv) : cadd(in.pwk, crmul(gam, in.hv));
    if (!is00) {
      vwk = cadd(vwk, crmul(fac, cadd(crmul(-k2, in.vor), x[0])));
      pwk = cadd(pwk, crmul(fac, cadd(crmul(-k2, in.phi), x[1])));
    }
    vwk = cadd(vwk, crmul(rdt, in.vor));
    pwk = cadd(pwk, crmul(rdt, in.phi));
    vwk = crmul(facb, vwk);
    pwk = crmul(facb, pwk);
    const double yj = (double)j * dy - 0.5 * Ly;
    const cplx m2 = mframe(kx, yj, f2);
    gst(vorwk + o, cmul(vwk, m2));
    gst(phiwk + o, cmul(pwk, m2));
  }
};

struct Implicit00Sink {
  typedef R2 In;
  cplx *u, *wk;
  const double* rf;
  double fac, facb, rdt, gam;
  int rk;
  __device__ In load(int j) const { return In{gld(u + j).x, rk == 1 ? 0.0 : gld(wk + j).x}; }
  __device__ void apply(int j, const In& in, const cplx* x) {
    const double rfj = gldd(rf + j);
    double wkj = (rk == 1) ? gam * rfj : in.wk + gam * rfj;
    wkj = wkj + fac * (-0.0 * in.u + x[0].x);
    wkj = facb * (wkj + rdt * in.u);
    gst(wk + j, mkc(wkj, 0.0));
  }
};

__global__ void __launch_bounds__(128)
k_y_implicit(YArgs a, int rk) {
  const long L = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long P = a.g.P;
  if (L >= P + 2) return;
  if (a.sc->halted) return;
  const int ir = rk - 1;
  double dt, f1, f2;
  substep_factors(a.sc, a.g.s, rk, dt, f1, f2);
  const double fac = c_alpha[ir] * (1.0 / a.g.Re);
  const double facb = -a.g.Re * c_ibeta[ir];
  const double rdt = 1.0 / dt;
  const double dtr1 = a.sc->dtr * c_ibeta[ir];
  const int N = a.g.my;
  if (L < P) {
    int i, k;
    double kx, k2;
    line_geom(a.g, L, i, k, kx, k2);
    const long LB = sline(a.g, i, k), FS = SNF_CW;   // the line in the fields
    {
      const DOp o = make_dop(a.b1, a.tw_old + i, gld(a.shp_old + i), N);
      const cplx* in[1] = {a.rdome + LB};
      cplx* ys[1] = {a.s1 + L};
      deriv_forward<1, true, 4>(o, a.c.rh1, in, FS, ys, a.PS, gld(a.shm_old + i));
      SubSink sk{a.rhv + LB, FS};
      deriv_backward<1, 4>(o, ys, a.PS, sk);
    }
    {
      const DOp o = make_dop(a.b2, a.tw_old + a.g.kxn + i, gld(a.shp_old + i), N);
      const cplx* in[2] = {a.vor + LB, a.phi + LB};
      cplx* ys[2] = {a.s1 + L, a.s2 + L};
      deriv_forward<2, false, 4>(o, a.c.rh2, in, FS, ys, a.PS, gld(a.shm_old + i));
      ImplicitRKSink sk{a.rhg + LB, a.rhv + LB, a.vor + LB, a.phi + LB, a.vorwk + LB, a.phiwk + LB,
                        FS, fac, facb, rdt, k2, kx, c_gama[ir], f2, a.g.dy, a.g.Ly, rk, L == 0};
      deriv_backward<2, 2>(o, ys, a.PS, sk);
    }
    // (8) Helmholtz solves for vor and phi with k2 + dtr1      (cross.f90:898-928)
    if (L != 0) {
      double lm[7];
      lapmat(a.c, dtr1 + k2, lm);
      HScratch hs{a.hs.hr + L, a.hs.hc + L};
      HTail tail;
      const cplx* in[2] = {a.vorwk + LB, a.phiwk + LB};
      cplx* ys[2] = {a.s1 + L, a.s2 + L};
      helm_forward<2, 4>(lm, a.c.pr2, N, gld(a.shp_new + i), gld(a.shm_new + i), in, FS, ys, a.PS, hs, tail);
      cplx* outs[2] = {a.vor + LB, a.phi + LB};
      StoreSink<2> sk{outs, FS};
      helm_backward<2, 4>(N, ys, a.PS, hs, tail, sk);
    }
    // (9) zeta * [hg, hv] -> [vorwk, phiwk] with the phase map (cross.f90:931-944)
    if (rk != 3) {
      const double zet = c_zeta[ir];
      for (int j = 0; j < N; ++j) {
        const long o = LB + (long)j * FS;
        const double yj = (double)j * a.g.dy - 0.5 * a.g.Ly;
        const cplx m1 = mframe(kx, yj, f1);
        gst(a.vorwk + o, cmul(crmul(zet, gld(a.rhg + o)), m1));
        gst(a.phiwk + o, cmul(crmul(zet, gld(a.rhv + o)), m1));
      }
    }
    if (L != 0) helm_line(a, L, k2, gld(a.shp_new + i), gld(a.shm_new + i), a.phi + LB, FS, a.s1 + L, a.v + LB, FS);
    const DOp o1n = make_dop(a.b1, a.tw_new + i, gld(a.shp_new + i), N);
    d1_one<8>(a, o1n, gld(a.shm_new + i), a.vor + LB, FS, a.s1 + L, a.dome + LB, FS);
    d1_one<8>(a, o1n, gld(a.shm_new + i), a.v + LB, FS, a.s2 + L, a.dv + LB, FS);
  } else {
    cplx* u = (L == P) ? a.u00c : a.w00c;
    cplx* wk = (L == P) ? a.u00wk : a.w00wk;
    const double* rf = (L == P) ? a.rf0u : a.rf0w;
    double* dd = (L == P) ? a.dudy : a.dwdy;
    {
      const DOp o = make_dop(a.b2, a.id2, mkc(1.0, 0.0), N);
      const cplx* in[1] = {u};
      cplx* ys[1] = {a.s1 + L};
      deriv_forward<1, false, 4>(o, a.c.rh2, in, 1, ys, a.PS, mkc(1.0, 0.0));
      Implicit00Sink sk{u, wk, rf, fac, facb, rdt, c_gama[ir], rk};
      deriv_backward<1, 4>(o, ys, a.PS, sk);
    }
    helm_line(a, L, dtr1, mkc(1.0, 0.0), mkc(1.0, 0.0), wk, 1, a.s1 + L, u, 1);
    if (rk != 3)
      for (int j = 0; j < N; ++j) gst(wk + j, mkc(c_zeta[ir] * gldd(rf + j), 0.0));
    d1_real(a, L, u, dd);
  }
}

// ---------------------------------------------------------------------------
// operator entry points (parity tests): nl independent lines, layout [y][l]
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(128)
k_op_deriv(int N, long nl, YCoef c, BaseFactor b, const TwistTail* tw,
                           const cplx* shp, const cplx* shm, const cplx* u, cplx* du, cplx* ys,
                           int order) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nl) return;
  const DOp o = make_dop(b, tw + l, gld(shp + l), N);
  const cplx* in[1] = {u + l};
  cplx* y[1] = {ys + l};
  if (order == 1)
    deriv_forward<1, true, 4>(o, c.rh1, in, nl, y, nl, gld(shm + l));
  else
    deriv_forward<1, false, 4>(o, c.rh2, in, nl, y, nl, gld(shm + l));
  cplx* out[1] = {du + l};
  StoreSink<1> sk{out, nl};
  deriv_backward<1, 4>(o, y, nl, sk);
}

__global__ void __launch_bounds__(128)
k_op_helm(int N, long nl, YCoef c, const double* k2, const cplx* shp,
                          const cplx* shm, const cplx* u, cplx* p, cplx* ys, HScratch hs) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nl) return;
  double lm[7];
  lapmat(c, gldd(k2 + l), lm);
  HScratch h{hs.hr + l, hs.hc + l};
  HTail tail;
  const cplx* in[1] = {u + l};
  cplx* y[1] = {ys + l};
  helm_forward<1, 4>(lm, c.pr2, N, gld(shp + l), gld(shm + l), in, nl, y, nl, h, tail);
  cplx* out[1] = {p + l};
  StoreSink<1> sk{out, nl};
  helm_backward<1, 4>(N, y, nl, h, tail, sk);
}

// ---------------------------------------------------------------------------
// energy (operators.f90:870-924): sums of |vor|^2 and |phi|^2 over every
// (y, kx, kz) with kx > 0 counted twice, and of U^2, W^2; a fixed-order
// two-pass reduction (per-block partials, then one block).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_energy_part(Geom g, const cplx* vor, const cplx* phi, double* part) {
  // element e of the kz-chunk layout: kz chunk t / my of kx i, kz % SNF_CW = e % SNF_CW
  const long n = sfield_len(g), chunk = (long)g.nkzb * g.my * SNF_CW;
  double sa = 0.0, sb = 0.0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long r = e % chunk;   // within kx = e / chunk
    const int k = (int)(r / ((long)g.my * SNF_CW)) * SNF_CW + (int)(e & (SNF_CW - 1));
    if (k >= g.mz) continue;
    const double wt = (e < chunk) ? 1.0 : 2.0;   // kx = 0 counted once
    sa += wt * cabs2(gld(vor + e));
    sb += wt * cabs2(gld(phi + e));
  }
  __shared__ double ra[256], rb[256];
  ra[threadIdx.x] = sa;
  rb[threadIdx.x] = sb;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      ra[threadIdx.x] += ra[threadIdx.x + st];
      rb[threadIdx.x] += rb[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    gstd(part + 2 * blockIdx.x, ra[0]);
    gstd(part + 2 * blockIdx.x + 1, rb[0]);
  }
}

__global__ void __launch_bounds__(256)
k_energy_final(const double* part, int nb, const cplx* u00, const cplx* w00, int my, double* out) {
  __shared__ double r[4][256];
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    s[0] += gldd(part + 2 * b);
    s[1] += gldd(part + 2 * b + 1);
  }
  for (int j = threadIdx.x; j < my; j += blockDim.x) {
    const double u = gld(u00 + j).x, w = gld(w00 + j).x;
    s[2] += u * u;
    s[3] += w * w;
  }
  for (int q = 0; q < 4; ++q) r[q][threadIdx.x] = s[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int q = 0; q < 4; ++q) r[q][threadIdx.x] += r[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 4) gstd(out + threadIdx.x, r[threadIdx.x][0]);
}

// phi = d2v - k2 v on every line but the (0,0) one (getv iopt=-1, visc3d
// with scal = 1, operators.f90:773-790)
__global__ void __launch_bounds__(256)
k_lap_combine(Geom g, const cplx* v, const cplx* d2v, cplx* phi) {
  const long n = (long)g.my * g.P;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long L = e % g.P;
    if (L == 0) continue;
    const int i = (int)(L / g.mz), k = (int)(L - (long)i * g.mz);
    const double kx = g.alp * (double)i, kz = kz_of(k, g.nz, g.mz, g.gam);
    const double k2 = kx * kx + kz * kz;
    const cplx a = gld(d2v + e), b = gld(v + e);
    gst(phi + e, mkc(a.x - k2 * b.x, a.y - k2 * b.y));
  }
}

// multiplier tables of the spectral y stage (snf_yfk.h)
__global__ void __launch_bounds__(256)
k_ysym(int N, YCoef c, const cplx* __restrict__ shp, const cplx* __restrict__ rt,
       cplx* __restrict__ out, const DevScalars* sc) {
  k_ysym_body(N, c, shp, rt, out, sc);
}
