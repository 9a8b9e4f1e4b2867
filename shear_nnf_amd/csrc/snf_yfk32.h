// The explicit y-stage kernel with its line transforms in FP32 (cross.f90:
// 565-796, the same substep as k_yf_explicit), an option of the mixed
// precision mode (SNF_Y32=1; the default mixed mode keeps FP64 transforms).
//
// In the mixed mode the x-z stage runs in FP32 and the six inter-stage
// fields (rhg, rhv, rdome in; v, dv, dome out) are stored FP32; this kernel
// also runs the transforms that produce them, and the viscous terms added to
// the FP32 right-hand sides, in FP32, and stores the RK accumulators
// (right-hand sides kept within a step) as FP32.  What stays FP64: the state
// vor, phi, U, W and its update vn = v + dt (gamma h + acc), the phase
// tables and the moving-frame maps applied to the state, the transforms of
// the 00 lines (the mean flow U, W).  The transforms run on FP32 columns
// against an FP32 copy of the plan's table: half the LDS bytes of the FP64
// transforms and fewer registers.  The price is accuracy: the rounding of a
// transform scales with the norm of its line, not with each coefficient, so
// small correlations such as the Reynolds stress move further from FP64.
//
// LDS of a block: the FP64 transfer tile [N][YF_LINES] (the state passes
// through it unrounded), two FP32 columns per wave, the FP32 table, the four
// FP64 phase tables, the FP64 multiplier rows where two blocks per CU still
// fit (else they are read from the global table).
// Radix plans only (PL::f32).
#pragma once
#include "snf_yfk.h"

__device__ __forceinline__ cplxf y32_f(cplx a) { return make_float2((float)a.x, (float)a.y); }

template <class PL>
struct Y32 {
  static constexpr int N = PL::NC;
  static constexpr int RTE = (PL::tw_len(PL::NC) + 1) & ~1;   // FP32 table entries (even)
  static constexpr size_t base = (size_t)N * YF_LINES * sizeof(cplx) +
                                 (size_t)2 * YF_LINES * N * sizeof(cplxf) +
                                 (size_t)RTE * sizeof(cplxf) +
                                 (size_t)4 * YPH_N(PL::Q) * sizeof(cplx);
  // the two multiplier rows (FP64, [2][N] each) staged in LDS where two
  // blocks still fit a CU (the kernel runs at 2 waves per SIMD), else read
  // from the global table
  static constexpr size_t rows = (size_t)4 * N * sizeof(cplx);
  static constexpr bool rows_lds = 2 * (base + rows) <= 160 * 1024;
  static constexpr size_t lds_bytes = base + (rows_lds ? rows : 0);
};
// row of kx i of the multiplier table tab, staged into dst (2 N entries)
// or read in place
template <bool LDS>
__device__ __forceinline__ YSymRow<!LDS> y32_row(cplx* dst, const cplx* tab, int i, int N) {
  if constexpr (!LDS) {
    return YSymRow<true>{tab + (long)i * N * YS, N};
  } else {
    for (int t = threadIdx.x; t < 2 * N; t += blockDim.x) {
      const long src = ((long)i * N + (t >> 1)) * YS + (t & 1);
      dst[(t & 1) * N + (t >> 1)] = gld(tab + src);
    }
    return YSymRow<false>{dst, N};
  }
}

template <class PL>
__device__ __forceinline__ void k_yf_explicit32_body(YArgs a, int rk) {
  static_assert(PL::f32 && !PL::direct, "FP32 transforms: radix plans");
  constexpr int Q = PL::Q, N = PL::NC;
  if (a.sc->halted) return;
  extern __shared__ cplx T[];
  const Geom& g = a.g;
  const long P = SNF_CW;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nch = (g.mz + YF_LINES - 1) / YF_LINES;
  const long nbl = (long)g.kxn * nch;
  const long lb = xcd_block(blockIdx.x, nbl);
  const int i = (int)(lb / nch), ch = (int)(lb - (long)i * nch);
  const long L0 = sline(g, i, YF_LINES * ch);
  const int nv = min(YF_LINES, g.mz - YF_LINES * ch);
  cplx pv[Q];
  if (blockIdx.x != nbl) yf_issue(pv, a.vor, P, L0, nv, N);
  cplxf* cols = reinterpret_cast<cplxf*>(T + YF_LINES * N);
  cplxf* A = cols + 2 * w * N;   // this wave's two FP32 columns
  cplxf* B = A + N;
  cplxf* rt = cols + 2 * YF_LINES * N;
  for (int t = threadIdx.x; t < PL::tw_len(N); t += blockDim.x) rt[t] = y32_f(gld(a.tpy + t));
  cplx* pho = reinterpret_cast<cplx*>(rt + Y32<PL>::RTE);
  __syncthreads();
  const int ir = rk - 1;
  double dt, f1, f2;
  substep_factors(a.sc, g.s, rk, dt, f1, f2);
  const double fac = 1.0 / g.Re;
  const double gam = c_gama[ir], zet = c_zeta[ir];
  const double invN = 1.0 / (double)N;
  const TileCol tc{w};
  const OwnCol oc{0};
  cplx x[Q];

  if (blockIdx.x == nbl) {
    // ---- 00 lines: U (wave 0) and W (wave 1), real, untwisted, in FP64
    // (the mean flow keeps FP64 transforms): two FP64 columns and the FP64
    // table in the transfer tile
    static_assert(PL::tw_len(N) <= 2 * N, "FP64 table of the 00 lines fits the tile");
    cplx* rt64 = T + 2 * N;
    for (int t = threadIdx.x; t < PL::tw_len(N); t += blockDim.x) rt64[t] = gld(a.tpy + t);
    __syncthreads();
    if (w >= 2) return;
    cplx* U = T + w * N;
    cplx* u = w == 0 ? a.u00c : a.w00c;
    cplx* uwk = w == 0 ? a.u00wk : a.w00wk;
    const double* rf = w == 0 ? a.rf0u : a.rf0w;
    double* dd = w == 0 ? a.dudy : a.dwdy;
    const cplx* sy = a.ysym_new;   // kx = 0 row: untwisted
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) {
        x[q] = gld(u + j);
        U[j] = x[q];
      }
    }
    PL::template fft_wave<-1>(U, oc, rt64, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) U[m] = cmul(U[m], gld(sy + YS * m + 1));
    }
    PL::template fft_wave<1>(U, oc, rt64, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) {
        const double rfj = gldd(rf + j) + fac * (U[j].x * invN);
        double wkj = (rk == 1) ? gam * rfj : gld(uwk + j).x + gam * rfj;
        const double un = x[q].x + dt * wkj;
        if (rk != 3) wkj = rfj * zet;
        gst(u + j, mkc(un, 0.0));
        gst(uwk + j, mkc(wkj, 0.0));
        U[j] = mkc(un, 0.0);
      }
    }
    PL::template fft_wave<-1>(U, oc, rt64, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) U[m] = cmul(U[m], gld(sy + YS * m));
    }
    PL::template fft_wave<1>(U, oc, rt64, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) gstd(dd + j, U[j].x * invN);
    }
    return;
  }

  // the RK accumulators hold FP32-accurate right-hand sides and live for
  // one step: stored FP32, the block's chunk at cplxf offset L0
  cplx* vwk_blk = reinterpret_cast<cplx*>(reinterpret_cast<cplxf*>(a.vorwk) + L0);
  cplx* pwk_blk = reinterpret_cast<cplx*>(reinterpret_cast<cplxf*>(a.phiwk) + L0);
  const int k = min(YF_LINES * ch + w, g.mz - 1);
  const double kx = g.alp * (double)i, kzv = kz_of(k, g.nz, g.mz, g.gam);
  const double k2 = kx * kx + kzv * kzv;
  const bool is00 = (i == 0 && k == 0);
  const cplx so = gld(a.shp_old + i), sn = gld(a.shp_new + i);
  const double tho = atan2(so.y, so.x) * invN, thn = atan2(sn.y, sn.x) * invN;
  cplx* phn = pho + YPH_N(Q);
  cplx* pm1 = phn + YPH_N(Q);
  cplx* pm2 = pm1 + YPH_N(Q);
  cplx* rows = pm2 + YPH_N(Q);
  const auto syo = y32_row<Y32<PL>::rows_lds>(rows, a.ysym_old, i, N);
  const auto syn = y32_row<Y32<PL>::rows_lds>(rows + 2 * N, a.ysym_new, i, N);
  {
    const double pa[4] = {tho, thn, kx * f1 * g.dy, kx * f2 * g.dy};
    const double pa0[4] = {0.0, 0.0, -0.5 * kx * f1 * g.Ly, -0.5 * kx * f2 * g.Ly};
    for (int t = threadIdx.x; t < 4 * YPH_N(Q); t += blockDim.x) pho[t] = yph_entry(t, Q, pa, pa0);
  }

  // ===== omega_y chain =====
  cplx pfa[Q];
  YPf<Q, YPfLive<PL>::value> pfb;
  yf_commit(T, pv, N);                     // vor
  yf_issue_io(pfa, a.rhg, P, L0, nv, N, 1);
  if (rk > 1) pfb.issue_blk_io(vwk_blk, N, 1);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        x[q] = T[tc(j)];
        A[j] = y32_f(cmul(x[q], e));
      }
    }
  }
  PL::template fft<-1>(A, oc, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s2 = syo.d2(m);
      const cplx mul = is00 ? mkc(0.0, 0.0) : crmul(fac * invN, mkc(s2.x - k2, s2.y));
      A[m] = wf_mul(A[m], y32_f(mul));
    }
  }
  PL::template fft<1>(A, oc, rt, N);
  yf_commit_io(T, pfa, N, 1);              // rhg
  yf_issue_io(pfa, a.rdome, P, L0, nv, N, 1);
  {
    YPh<false> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) A[j] = y32_f(cadd(T[tc(j)], cmul(to_d(A[j]), e)));   // hg
    }
  }
  if (rk > 1) pfb.commit_io(T, N, 1);      // vorwk
  pfb.issue(a.phi, P, L0, nv, N);
  {
    YPh<false> m1(pm1, lane);
    YPh<false> m2(pm2, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx p1 = m1.at(q), p2 = m2.at(q);
      if (j < N) {
        const cplx hgj = to_d(A[j]);
        cplx vwk = (rk == 1) ? crmul(gam, hgj) : cadd(T[tc(j)], crmul(gam, hgj));
        const cplx vn = cadd(x[q], crmul(dt, vwk));
        if (rk != 3) vwk = cmul(crmul(zet, hgj), p1);
        x[q] = cmul(vn, p2);
        T[tc(j)] = vwk;
      }
    }
  }
  // the third substep's accumulators are never read (a step starts afresh)
  if (rk != 3) yf_flush_blk_io(T, vwk_blk, N, 1);
  {
    YPh<true> ph(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        T[tc(j)] = x[q];
        A[j] = y32_f(cmul(x[q], e));
      }
    }
  }
  yf_flush(T, a.vor, P, L0, nv, N);
  PL::template fft<-1>(A, oc, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) A[m] = wf_mul(A[m], y32_f(crmul(invN, syn.d1(m))));
  }
  PL::template fft<1>(A, oc, rt, N);
  {
    YPh<false> ph(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[tc(j)] = cmul(to_d(A[j]), e);
    }
  }
  yf_flush_io(T, a.dome, P, L0, nv, N, 1);

  // ===== phi chain =====
  yf_commit_io(T, pfa, N, 1);              // rdome
  yf_issue_io(pfa, a.rhv, P, L0, nv, N, 1);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) A[j] = y32_f(cmul(T[tc(j)], e));
    }
  }
  PL::template fft<-1>(A, oc, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) A[m] = wf_mul(A[m], y32_f(crmul(-invN, syo.d1(m))));
  }
  pfb.commit(T, N);                        // phi
  if (rk > 1) pfb.issue_blk_io(pwk_blk, N, 1);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        x[q] = T[tc(j)];
        B[j] = y32_f(cmul(x[q], e));
      }
    }
  }
  PL::template fft<-1>(B, oc, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s2 = syo.d2(m);
      const cplx mul = is00 ? mkc(0.0, 0.0) : crmul(fac * invN, mkc(s2.x - k2, s2.y));
      A[m] = wf_add(A[m], wf_mul(B[m], y32_f(mul)));
    }
  }
  PL::template fft<1>(A, oc, rt, N);
  yf_commit_io(T, pfa, N, 1);              // rhv
  {
    YPh<false> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) A[j] = y32_f(cadd(T[tc(j)], cmul(to_d(A[j]), e)));   // hv
    }
  }
  if (rk > 1) pfb.commit_io(T, N, 1);      // phiwk
  {
    YPh<false> m1(pm1, lane);
    YPh<false> m2(pm2, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx p1 = m1.at(q), p2 = m2.at(q);
      if (j < N) {
        const cplx hvj = to_d(A[j]);
        cplx pwk = (rk == 1) ? crmul(gam, hvj) : cadd(T[tc(j)], crmul(gam, hvj));
        const cplx pn = cadd(x[q], crmul(dt, pwk));
        if (rk != 3) pwk = cmul(crmul(zet, hvj), p1);
        x[q] = cmul(pn, p2);
        T[tc(j)] = pwk;
      }
    }
  }
  if (rk != 3) yf_flush_blk_io(T, pwk_blk, N, 1);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = lane + 64 * q;
    if (j < N) T[tc(j)] = x[q];
  }
  yf_flush(T, a.phi, P, L0, nv, N);
  // v = lapcdy(phi) and dv = D1(v) with the new twist; the (0,0) mode keeps
  // its v (lapcdy is not called for it) and only takes dv = D1(v)
  if (L0 == 0) yf_fill_io(T, a.v, P, L0, nv, N, 1);   // block of the (0,0) mode
  {
    YPh<true> ph(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        if (is00) x[q] = T[tc(j)];             // v of the (0,0) mode, unchanged
        A[j] = y32_f(cmul(x[q], e));
      }
    }
  }
  PL::template fft<-1>(A, oc, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s1 = syn.d1(m);
      if (!is00) {
        const cplx s2 = syn.d2(m);
        const cplxf H = y32_f(yf_div(mkc(invN, 0.0), mkc(s2.x - k2, s2.y)));   // lapcdy
        const cplxf vh = wf_mul(A[m], H);
        B[m] = wf_mul(vh, y32_f(s1));
        A[m] = vh;
      } else {
        A[m] = wf_mul(A[m], y32_f(crmul(invN, s1)));
      }
    }
  }
  PL::template fft<1>(A, oc, rt, N);
  PL::template fft<1>(B, oc, rt, N, !is00);
  {
    YPh<false> ph(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[tc(j)] = cmul(to_d(is00 ? A[j] : B[j]), e);   // dv
    }
  }
  yf_flush_io(T, a.dv, P, L0, nv, N, 1);
  {
    YPh<false> ph(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[tc(j)] = is00 ? x[q] : cmul(to_d(A[j]), e);   // v
    }
  }
  yf_flush_io(T, a.v, P, L0, nv, N, 1);
}

template <class PL>
__global__ void __launch_bounds__(64 * YF_LINES) __attribute__((amdgpu_waves_per_eu(2)))
k_yf_explicit32(YArgs a, int rk) {
  k_yf_explicit32_body<PL>(a, rk);
}
template <class PL>
__global__ void __launch_bounds__(64 * YF_LINES) __attribute__((amdgpu_waves_per_eu(2)))
k_yf_explicit32_tab(const YArgs* __restrict__ t, int rk) {
  k_yf_explicit32_body<PL>(t[blockIdx.z], rk);
}
