// Scalar and per-kx kernels of one RK3 substep: twisted border tails of the
// derivative factors, the CFL time step and statistics record, the shear
// offsets, and step/call control (cross.f90:81-103, 353-406, 1598-1712,
// 2067-2119).
#include "snf_internal.h"
#include "snf_lines.h"
#include "snf_rk.h"



// exp(i kx y_j m) of moving_frame (operators.f90:86-145)
__device__ __forceinline__ cplx mframe(double kx, double yj, double m) {
  double sn, cs;
  sincos((kx * yj) * m, &sn, &cs);
  return mkc(cs, sn);
}

// ---------------------------------------------------------------------------
// per-kx twisted border tail of a derivative factor (modChol_addshear with
// the border pre-multiplication of derivyc); one wave per (kx, operator)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// istop = 1 with the condition that raised it (snf_stop_reason), the first
// one of the run kept
__device__ __forceinline__ void raise_stop(DevScalars* sc, int reason) {
  if (!sc->istop) sc->stop_reason = reason;
  sc->istop = 1;
}

__global__ void __launch_bounds__(64)
k_twist(int N, BaseFactor b1, BaseFactor b2, YCoef c, const cplx* __restrict__ shp_arr,
        TwistTail* __restrict__ out, int nk, const DevScalars* sc) {
  if (sc && sc->halted) return;
  const int kx = blockIdx.x, op = blockIdx.y, lane = threadIdx.x;
  const BaseFactor b = op == 0 ? b1 : b2;
  const double* a = op == 0 ? c.pr1 : c.pr2;
  const double a0 = a[3], a1 = a[4], a2 = a[5], a3 = a[6];
  const cplx t = shp_arr[kx];
  __shared__ cplx sLe[3], sLf[3], sLg[3];
  auto Ler = [&](int k) { return crmul(b.Ler[k], t); };
  auto Lfr = [&](int k) { return crmul(b.Lfr[k], t); };
  auto Lgr = [&](int k) { return crmul(b.Lgr[k], t); };
  auto rec = [&](cplx x3, cplx x2, cplx x1, int i, double av) {
    cplx r;
    r.x = b.Dm1[i] * (av - x3.x * b.D[i - 3] * b.Lc[i] - x2.x * b.D[i - 2] * b.Lb[i] - x1.x * b.D[i - 1] * b.La[i]);
    r.y = b.Dm1[i] * (0.0 - x3.y * b.D[i - 3] * b.Lc[i] - x2.y * b.D[i - 2] * b.Lb[i] - x1.y * b.D[i - 1] * b.La[i]);
    return r;
  };
  if (lane == 0) {
    const cplx e0 = rec(Ler(N - 9), Ler(N - 8), Ler(N - 7), N - 6, a3);
    const cplx e1 = rec(Ler(N - 8), Ler(N - 7), e0, N - 5, a2);
    const cplx e2 = rec(Ler(N - 7), e0, e1, N - 4, a1);
    sLe[0] = e0; sLe[1] = e1; sLe[2] = e2;
    const cplx f0 = rec(Lfr(N - 8), Lfr(N - 7), Lfr(N - 6), N - 5, a3);
    const cplx f1 = rec(Lfr(N - 7), Lfr(N - 6), f0, N - 4, a2);
    sLf[0] = f0; sLf[1] = f1;
    sLg[0] = rec(Lgr(N - 7), Lgr(N - 6), Lgr(N - 5), N - 4, a3);
  }
  __syncthreads();
  // tail entry q of a 3-row block held in LDS
  auto sel3 = [](const cplx* t, int q) { return t[q]; };
  auto Le = [&](int k) { return k < N - 6 ? Ler(k) : sel3(sLe, k - (N - 6)); };
  auto Lf = [&](int k) { return k < N - 5 ? Lfr(k) : sel3(sLf, k - (N - 5)); };
  auto Lg = [&](int k) { return k < N - 4 ? Lgr(k) : sel3(sLg, k - (N - 4)); };
  // D(N-3) = a0 - sum_{k<=N-4} |Le|^2 D
  double s_ee = 0.0;
  cplx s_fe = mkc(0, 0), s_ge = mkc(0, 0);
  for (int k = lane; k <= N - 4; k += 64) {
    const cplx le = Le(k), lf = Lf(k), lg = Lg(k);
    s_ee += cabs2(le) * b.D[k];
    s_fe = cadd(s_fe, crmul(b.D[k], cmulc(lf, le)));
    s_ge = cadd(s_ge, crmul(b.D[k], cmulc(lg, le)));
  }
  s_ee = wave_sum(s_ee);
  s_fe.x = wave_sum(s_fe.x); s_fe.y = wave_sum(s_fe.y);
  s_ge.x = wave_sum(s_ge.x); s_ge.y = wave_sum(s_ge.y);
  const double DN3 = a0 - s_ee;
  const double dN3 = 1.0 / DN3;
  const cplx lfN3 = crmul(dN3, csub(mkc(a1, 0.0), s_fe));
  const cplx lgN3 = crmul(dN3, csub(mkc(a2, 0.0), s_ge));
  if (lane == 0) { sLf[2] = lfN3; sLg[1] = lgN3; }
  __syncthreads();
  double s_ff = 0.0;
  cplx s_gf = mkc(0, 0);
  for (int k = lane; k <= N - 3; k += 64) {
    const double Dk = (k == N - 3) ? DN3 : b.D[k];
    const cplx lf = Lf(k), lg = Lg(k);
    s_ff += cabs2(lf) * Dk;
    s_gf = cadd(s_gf, crmul(Dk, cmulc(lg, lf)));
  }
  s_ff = wave_sum(s_ff);
  s_gf.x = wave_sum(s_gf.x); s_gf.y = wave_sum(s_gf.y);
  const double DN2 = a0 - s_ff;
  const double dN2 = 1.0 / DN2;
  const cplx lgN2 = crmul(dN2, csub(mkc(a1, 0.0), s_gf));
  if (lane == 0) sLg[2] = lgN2;
  __syncthreads();
  double s_gg = 0.0;
  for (int k = lane; k <= N - 2; k += 64) {
    const double Dk = (k == N - 3) ? DN3 : ((k == N - 2) ? DN2 : b.D[k]);
    s_gg += cabs2(Lg(k)) * Dk;
  }
  s_gg = wave_sum(s_gg);
  const double DN1 = a0 - s_gg;
  if (lane == 0) {
    TwistTail tt;
    for (int q = 0; q < 3; ++q) { tt.Le[q] = sLe[q]; tt.Lf[q] = sLf[q]; tt.Lg[q] = sLg[q]; }
    tt.Dm1[0] = dN3;
    tt.Dm1[1] = dN2;
    tt.Dm1[2] = 1.0 / DN1;
    out[op * nk + kx] = tt;
  }
}


// ---------------------------------------------------------------------------
// scalar kernel: statistics accumulation, CFL time step, UNER record,
// shear-offset update (cross.f90:1598-1712, 2067-2119, 726-746)
// ---------------------------------------------------------------------------
struct ScalarArgs {
  Geom g;
  DevScalars* sc;
  int nhist, readflag, explicit_visc;
  double cflx, cfly, cflz, cflvv, CFL, uprim;
  unsigned long long* umax_acc;
  unsigned long long* upart;  // [nupart][4] maxima as double bit patterns (or null)
  long nupart;
  const double* stat_plane;   // [my][STAT_W]
  const double* pplane;       // [my][3]: 'p' partials summed per plane (k_phist_plane)
  double *stats, *stats2;     // [my][NSTAT]
  double* ring;               // [cap][24]
  int cap;
  cplx *shp_new, *shm_new;
  double wrap_lx;             // Lx when the offsets wrap (iydealiasing=2), else 0
};

// The 'p' triple products of a statistics step (cross.f90:2030-2045): block
// j sums the x-row kernel's partials of plane j in a fixed order (lanes in
// strides of 64, then a shuffle tree).  Launched at rkstep 1 after the x-row
// kernel; a no-op off the statistics steps.
__device__ __forceinline__ void k_phist_plane_body(const double* __restrict__ phist, int nzb, double* __restrict__ pplane,
              const DevScalars* sc, int nhist) {
  if (sc->halted || (sc->steps_local - 1) % nhist != 0) return;
  const int j = blockIdx.x, lane = threadIdx.x;
  double ph[3] = {0.0, 0.0, 0.0};
  for (int zb = lane; zb < nzb; zb += 64)
    for (int q = 0; q < 3; ++q) ph[q] += phist[((long)j * nzb + zb) * 3 + q];
  for (int o = 32; o > 0; o >>= 1)
    for (int q = 0; q < 3; ++q) ph[q] += __shfl_xor(ph[q], o, 64);
  if (lane == 0)
    for (int q = 0; q < 3; ++q) pplane[(long)j * 3 + q] = ph[q];
}
__global__ void __launch_bounds__(64)
k_phist_plane(const double* __restrict__ phist, int nzb, double* __restrict__ pplane,
              const DevScalars* sc, int nhist) {
  k_phist_plane_body(phist, nzb, pplane, sc, nhist);
}

__device__ __forceinline__ void k_scalar_body(ScalarArgs a, int rk) {
  DevScalars* sc = a.sc;
  if (sc->halted) return;
  const Geom& g = a.g;
  const int tid = threadIdx.x;
  const int ihist = (rk == 1) && ((sc->steps_local - 1) % a.nhist == 0);
  __shared__ double sh_dt;
  const double dt_old = sc->Deltat;
  if (ihist) {
    // one thread per plane accumulates the plane's statistics
    for (int j = tid; j < g.my; j += blockDim.x) {
      const double* ph = a.pplane + (long)j * 3;
      const double* sp = a.stat_plane + (long)j * STAT_W;
      const double o3s = sp[16] - g.s;
      const double pk = -o3s * (-sp[3]) + o3s * o3s * (1.0 / g.Re);
      const double v[SNF_NSTAT] = {sp[11], sp[12], sp[13], sp[0], sp[1], sp[2], sp[3], sp[4],
                                   sp[5], sp[14], sp[15], sp[16], sp[6], sp[7], sp[8], sp[10],
                                   ph[0], ph[1], ph[2], pk, sp[9], 0.0, 0.0};
      double* st = a.stats + (long)j * SNF_NSTAT;
      double* st2 = a.stats2 + (long)j * SNF_NSTAT;
      for (int q = 0; q < 21; ++q) {
        st[q] += v[q];
        st2[q] += v[q] * dt_old;
      }
    }
  }
  // maxima of the per-row partials (order-free)
  __shared__ double red_um[4][256];
  if (rk == 1) {
    double lm[4] = {0.0, 0.0, 0.0, 0.0};
    // fold the accumulators and clear them for the next rk = 1 stage
    for (long n = tid; n < a.nupart; n += blockDim.x)
      for (int q = 0; q < 4; ++q) {
        lm[q] = fmax(lm[q], __longlong_as_double((long long)a.upart[n * 4 + q]));
        a.upart[n * 4 + q] = 0ull;
      }
    for (int q = 0; q < 4; ++q) red_um[q][tid] = lm[q];
  }
  __syncthreads();
  if (rk == 1) {
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
      if (tid < st)
        for (int q = 0; q < 4; ++q) red_um[q][tid] = fmax(red_um[q][tid], red_um[q][tid + st]);
      __syncthreads();
    }
  }
  if (tid == 0) {
    if (rk == 1) {
      double um[4];
      for (int q = 0; q < 4; ++q) {
        um[q] = fmax(__longlong_as_double((long long)a.umax_acc[q]), red_um[q][0]);
        a.umax_acc[q] = 0ull;
      }
      if (sc->steps_local == 1 && a.readflag == 0) {
        const double init = fabs(g.s) > 0.5 ? 1.0 : a.uprim;
        for (int q = 0; q < 4; ++q) um[q] = fmax(init, um[q]);
      }
      for (int q = 0; q < 4; ++q) sc->umax[q] = um[q];
      double dt = 1.0 / fmax(fmax(um[0] * a.cflx, um[2] * a.cfly), um[3] * a.cflz);
      sc->CFLv = a.CFL * a.cflvv * dt;
      if (sc->CFLv > a.CFL && a.explicit_visc == 1) dt = 1.0 / a.cflvv;
      double Deltat = dt;
      double dtr = g.Re / Deltat;
      if (Deltat < 1e-10) raise_stop(sc, 2);
      if (sc->time + Deltat > sc->dumptime) {
        Deltat = fabs(sc->dumptime - sc->time);
        dtr = g.Re / Deltat;
        if (sc->dumptime > sc->etime - 1e-13) raise_stop(sc, 1);
        sc->dumptime = sc->dumptime + sc->dtimag;
        sc->idump = 1;
      } else if (sc->time + Deltat > sc->etime) {
        Deltat = fabs(sc->etime - sc->time);
        dtr = g.Re / Deltat;
        raise_stop(sc, 1);
      }
      sc->Deltat = Deltat;
      sc->dtr = dtr;
    }
    if (ihist) {
      double ener[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, ener2[4] = {0, 0, 0, 0};
      for (int j = 0; j < g.my; ++j) {
        const double* sp = a.stat_plane + (long)j * STAT_W;
        for (int q = 0; q < 10; ++q) ener[q] += (q < 9) ? sp[q] : sp[10];
        const double o3s = sp[16] - g.s;
        ener2[0] += -o3s * (-sp[3]) + o3s * o3s * (1.0 / g.Re);
        ener2[1] += sp[9];
      }
      sc->nacum += 1;
      sc->total_dt += sc->Deltat;
      const double my = (double)g.my;
      double uner[10];
      for (int q = 0; q < 3; ++q) uner[q] = sqrt(fabs(ener[q] / my));
      for (int q = 3; q < 6; ++q) uner[q] = ener[q] / my;
      for (int q = 6; q < 9; ++q) uner[q] = sqrt(fabs(ener[q] / my));
      uner[9] = ener[9] / my;
      double* r = a.ring + (long)(sc->nrec % a.cap) * 24;
      r[0] = sc->time;
      r[1] = sc->Deltat;
      for (int q = 0; q < 10; ++q) r[2 + q] = uner[q];
      r[12] = a.CFL;
      r[13] = sc->CFLv;
      r[14] = sc->umax[1];
      r[15] = sc->umax[2];
      r[16] = sc->umax[3];
      r[17] = sc->xoffb;
      r[18] = sc->xofft;
      r[19] = sc->umax[0];
      for (int q = 0; q < 4; ++q) r[20 + q] = ener2[q] / my;
      sc->nrec += 1;
      if (isnan(uner[0])) raise_stop(sc, 3);
      if (uner[1] / (2.0 * M_PI / g.gam) < 1e-5) raise_stop(sc, 4);
    }
    sh_dt = sc->Deltat;
  }
  __syncthreads();
  // shear-periodic offsets for the next stage (cross.f90:726-746 / 866-884)
  const double Deltat = sh_dt;
  const double r1 = c_ca[rk - 1] * Deltat;
  const double thr = a.explicit_visc == 1 ? 0.01 : 0.5;
  double xwkb = sc->xwkb, xwkt = sc->xwkt;
  if (fabs(g.s) > thr) {
    xwkb = sc->xoffb + r1 * (g.s * (-g.Ly));
    xwkt = sc->xofft + r1 * (g.s * g.Ly);
    if (a.explicit_visc == 1 && a.wrap_lx > 0.0) {   // iydealiasing=2 (cross.f90:733-736)
      const double q = xwkb / a.wrap_lx;
      xwkb = (q - trunc(q)) * a.wrap_lx;
      xwkt = -xwkb;
    }
  }
  for (int i = tid; i < g.kxn; i += blockDim.x) {
    const double kx = g.alp * (double)i;
    double sn, cs;
    sincos(kx * xwkt, &sn, &cs);
    a.shp_new[i] = mkc(cs, -sn);
    sincos(kx * xwkb, &sn, &cs);
    a.shm_new[i] = mkc(cs, -sn);
  }
  __syncthreads();
  if (tid == 0) {
    sc->xwkb = xwkb;
    sc->xwkt = xwkt;
    if (rk == 3) {
      sc->xoffb = xwkb;
      sc->xofft = xwkt;
      sc->time = sc->time + Deltat;
      sc->steps_done += 1;
    }
  }
}
__global__ void __launch_bounds__(256)
k_scalar(ScalarArgs a, int rk) {
  k_scalar_body(a, rk);
}

// start of an advance call (cross.f90:81-103, 287-291, 383-406).  resume:
// the previous call stopped at a dump and this one continues the same run,
// as the reference's step loop does after writing the dump (cross.f90:356-373):
// the step count (nhist phase, the step-2 rule) carries on.
__global__ void k_begin_call(Geom g, DevScalars* sc, cplx* shp, cplx* shm, int resume) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    sc->halted = 0;
    if (!resume) {
      sc->steps_local = 0;
      sc->istop = 0;
      sc->stop_reason = 0;
    }
    sc->idump = 0;
    sc->xwkb = sc->xoffb;
    sc->xwkt = sc->xofft;
  }
  const double xb = sc->xoffb, xt = sc->xofft;
  for (int i = tid; i < g.kxn; i += blockDim.x) {
    const double kx = g.alp * (double)i;
    double sn, cs;
    sincos(kx * xt, &sn, &cs);
    shp[i] = mkc(cs, -sn);
    sincos(kx * xb, &sn, &cs);
    shm[i] = mkc(cs, -sn);
  }
}

// the reference checks istop/idump at the top of every step except the
// second one (cross.f90:353-374)
// wrap_lx > 0 (iydealiasing=2): the offsets are first wrapped into one
// period, xoffb = (xoffb/Lx - int(xoffb/Lx)) Lx, xofft = -xoffb (cross.f90:348-352)
__device__ __forceinline__ void k_step_begin_body(DevScalars* sc, double wrap_lx) {
  if (sc->halted) return;
  if (wrap_lx > 0.0) {
    const double q = sc->xoffb / wrap_lx;
    sc->xoffb = (q - trunc(q)) * wrap_lx;
    sc->xofft = -sc->xoffb;
    sc->xwkb = sc->xoffb;
    sc->xwkt = sc->xofft;
  }
  const int next = sc->steps_local + 1;
  if (next - 1 != 1 && (sc->istop || sc->idump)) {
    sc->halted = 1;
    return;
  }
  sc->steps_local = next;
}
__global__ void k_step_begin(DevScalars* sc, double wrap_lx) {
  k_step_begin_body(sc, wrap_lx);
}

// Layout change between the reference's jik order (y fastest, then kx, then
// kz: element ((k kxn + i) my + j)) and the device's kz-chunk order
// (sidx(g, j, i, k)), the transpose of chjik2ikj / chikj2jik
// (change.allfac_ieor.f:80-87, 117-124) on one GPU.  For each kx a block
// moves a 32 x 32 (y, kz) tile through LDS: the jik side is read or written
// in runs of 32 y, the device side in runs of SNF_CW kz over 32 consecutive y.
// dir = +1: jik -> device; dir = -1: the reverse.
#define JIK_T 32
__global__ void __launch_bounds__(256)
k_jik_transpose(Geom g, const cplx* __restrict__ src, cplx* __restrict__ dst, int dir) {
  __shared__ cplx t[JIK_T][JIK_T + 1];
  const int i = blockIdx.z;                       // kx
  const int tj = blockIdx.x * JIK_T, tk = blockIdx.y * JIK_T;
  const int tx = threadIdx.x & (JIK_T - 1), ty = threadIdx.x / JIK_T;   // 32 x 8
  const long kxn = g.kxn, my = g.my, mz = g.mz;
  if (dir > 0) {
    // read rows of j (contiguous in jik) for each k of the tile
    for (int r = ty; r < JIK_T; r += 8) {
      const int k = tk + r, j = tj + tx;
      if (k < mz && j < my) t[r][tx] = gld(src + ((long)k * kxn + i) * my + j);
    }
    __syncthreads();
    // write the device side: lanes in (kz % SNF_CW, y, kz / SNF_CW) order
    for (int r = ty; r < JIK_T; r += 8) {
      const int u = r * JIK_T + tx;
      const int kk = (u & (SNF_CW - 1)) + SNF_CW * (u / (SNF_CW * JIK_T));
      const int jj = (u / SNF_CW) % JIK_T;
      const int j = tj + jj, k = tk + kk;
      if (k < mz && j < my) gst(dst + sidx(g, j, i, k), t[kk][jj]);
    }
  } else {
    for (int r = ty; r < JIK_T; r += 8) {
      const int u = r * JIK_T + tx;
      const int kk = (u & (SNF_CW - 1)) + SNF_CW * (u / (SNF_CW * JIK_T));
      const int jj = (u / SNF_CW) % JIK_T;
      const int j = tj + jj, k = tk + kk;
      if (k < mz && j < my) t[kk][jj] = gld(src + sidx(g, j, i, k));
    }
    __syncthreads();
    for (int r = ty; r < JIK_T; r += 8) {
      const int k = tk + r, j = tj + tx;
      if (k < mz && j < my) gst(dst + ((long)k * kxn + i) * my + j, t[r][tx]);
    }
  }
}

// kx = 0 Hermitian fix on the state (cross.f90:287-291, set_conj)
__global__ void k_set_conj(Geom g, cplx* vor, cplx* phi) {
  const int j = blockIdx.x;
  for (int k = g.nz + 1 + threadIdx.x; k < g.mz; k += blockDim.x) {
    const long d = sidx(g, j, 0, k), s = sidx(g, j, 0, g.mz - k);
    vor[d] = cconj(vor[s]);
    phi[d] = cconj(phi[s]);
  }
}

// Layout change between [y][kx][kz] order (((j kxn + i) mz + k), the order of
// the host state vectors and operator arrays) and the device's kz-chunk
// order, for fields of element type E on the device side (FP64, or FP32 for
// the inter-stage fields of the mixed precision mode); the [y][kx][kz] side
// is always FP64.  One thread per element, in device order (dir = +1:
// [y][kx][kz] -> device; -1: the reverse).  The device side is read or
// written whole, the other side in runs of SNF_CW.
template <class E>
__global__ void __launch_bounds__(256) k_relayout(Geom g, const void* __restrict__ src,
                                                  void* __restrict__ dst, int dir) {
  const long n = sfield_len(g), chunk = (long)g.nkzb * g.my * SNF_CW;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / chunk);
    const long r = e - (long)i * chunk;
    const int kb = (int)(r / ((long)g.my * SNF_CW));
    const int j = (int)(r / SNF_CW - (long)kb * g.my);
    const int k = kb * SNF_CW + (int)(e & (SNF_CW - 1));
    if (k >= g.mz) continue;
    const long p = ((long)j * g.kxn + i) * g.mz + k;
    if (dir > 0)
      gst(reinterpret_cast<E*>(dst) + e, to_h<E>(gld(reinterpret_cast<const cplx*>(src) + p)));
    else
      gst(reinterpret_cast<cplx*>(dst) + p, to_d(gld(reinterpret_cast<const E*>(src) + e)));
  }
}

