// Explicit RK3 substep of the y stage (cross.f90:565-796) with the compact
// y-operators applied spectrally (snf_yfft.h): one wave per (kx,kz) line,
// YF_LINES lines per block.  The block moves every field through an LDS tile
// of [my][YF_LINES] (one contiguous run of the kz-chunk layout); each wave
// transforms its line in a private LDS column (and, for a second spectrum,
// in its column of the tile).
//
// Per line, with the old twist theta (shp at the start of the substep) and
// the new twist theta' (after the shift update):
//   omega_y chain (cross.f90:652-781):
//     hg  = rhg + nu (D2 - k2) vor           (visc3d; not for the (0,0) mode)
//     RK update of vor, vorwk, phase maps    -> vor, vorwk
//     dome = D1(vor)  with theta'            (cross.f90:434-445 of the next stage)
//   phi chain (cross.f90:607, 652-781):
//     hv  = rhv - D1(rdome) + nu (D2 - k2) phi
//     RK update of phi, phiwk, phase maps    -> phi, phiwk
//     v = lapcdy(phi), dv = D1(v) with theta' (cross.f90:791)
// Fields read: vor rhg vorwk rdome phi rhv phiwk; written: vor vorwk dome phi
// phiwk v dv -- each once.  The real x-z means U, W (00 lines, m=1 calls)
// take one extra block with one wave each.
#pragma once
#include "snf_yargs.h"
#include "snf_yfft.h"

#define YS 3   // multipliers per mode in the tables: d/dy, d2/dy2, filter

__device__ __forceinline__ cplx yf_expi(double a) {
  double s, c;
  sincos(a, &s, &c);
  return mkc(c, s);
}

// exp(i (a j + a0)) along the rows j = lane + 64 q of a lane: base * step^q
struct YPhase {
  cplx p, st;
  __device__ YPhase(double a, double a0, int lane) {
    p = yf_expi(a * (double)lane + a0);
    st = yf_expi(a * 64.0);
  }
  __device__ cplx next() {
    const cplx r = p;
    p = cmul(p, st);
    return r;
  }
};

// Phases exp(+-i (a j + a0)) of a block at the rows j = lane + 64 q, from a
// table of YPH_N(Q) values that the block fills once:
//   [exp(i a l), l < 8] [exp(i 8 a l), l < 8] [exp(i (64 a q + a0)), q < Q]
// so that a phase costs two table reads per lane and one product per row.
#define YPH_N(Q) (16 + (Q))
template <bool NEG>
struct YPh {
  cplx p;
  const cplx* c;
  __device__ YPh(const cplx* tab, int lane) : c(tab + 16) {
    p = cmul(tab[lane & 7], tab[8 + (lane >> 3)]);
  }
  __device__ cplx at(int q) const {
    const cplx e = cmul(p, c[q]);
    return NEG ? cconj(e) : e;
  }
};
// entry t of the tables of nb phases (a[b], a0[b]), filled by the block
__device__ __forceinline__ cplx yph_entry(int t, int Q, const double* a, const double* a0) {
  const int b = t / YPH_N(Q), r = t - b * YPH_N(Q);
  const double ang = r < 8 ? a[b] * (double)r
                   : (r < 16 ? a[b] * (double)(8 * (r - 8)) : a[b] * (double)(64 * (r - 16)) + a0[b]);
  return yf_expi(ang);
}

// The multiplier rows d/dy, d2/dy2 of a block's kx at one twist.  Staged in
// LDS as [2][m], or, for plans whose LDS would then admit one block per CU
// (PL::sym_gl), read from the global table [kx][m][YS], where the blocks of
// one kx (kept on one XCD) find them in L2.
template <bool GL> struct YSymRow;
template <> struct YSymRow<false> {
  const cplx* p;
  int N;
  __device__ __forceinline__ cplx d1(int m) const { return p[m]; }
  __device__ __forceinline__ cplx d2(int m) const { return p[N + m]; }
};
template <> struct YSymRow<true> {
  const cplx* p;
  int N;
  __device__ __forceinline__ cplx d1(int m) const { return gld(p + YS * m); }
  __device__ __forceinline__ cplx d2(int m) const { return gld(p + YS * m + 1); }
};
// row of kx i of the table tab: staged into the LDS slot dst (2 N entries,
// block-collective; read after the next block barrier) or pointed at in place
template <class PL>
__device__ __forceinline__ YSymRow<PL::sym_gl> ys_row(cplx* dst, const cplx* tab, int i, int N) {
  if constexpr (PL::sym_gl) {
    return YSymRow<true>{tab + (long)i * N * YS, N};
  } else {
    for (int t = threadIdx.x; t < 2 * N; t += blockDim.x) {
      const long src = ((long)i * N + (t >> 1)) * YS + (t & 1);
      dst[(t & 1) * N + (t >> 1)] = gld(tab + src);
    }
    return YSymRow<false>{dst, N};
  }
}
// LDS entries of one staged row
template <class PL> __host__ __device__ constexpr int ys_lds(int N) { return PL::sym_gl ? 0 : 2 * N; }

// block-collective transfers of field F (y stride P) for the nv lines from L0:
// element (j, c) at F + j P + L0 + c.  In the kz-chunk layout P = SNF_CW and
// the block's lines are one contiguous run.
static_assert(YF_LINES == SNF_CW, "a y block is one kz chunk");
__device__ __forceinline__ void yf_fill(cplx* T, const cplx* F, long P, long L0, int nv, int N) {
  __syncthreads();
  for (int e = threadIdx.x; e < N * YF_LINES; e += blockDim.x) {
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    T[TileCol{c}(j)] = (c < nv) ? gld(F + (long)j * P + L0 + c) : mkc(0.0, 0.0);
  }
  __syncthreads();
}
__device__ __forceinline__ void yf_flush(const cplx* T, cplx* F, long P, long L0, int nv, int N) {
  __syncthreads();
  for (int e = threadIdx.x; e < N * YF_LINES; e += blockDim.x) {
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    if (c < nv) gst(F + (long)j * P + L0 + c, T[TileCol{c}(j)]);
  }
  __syncthreads();
}

// The same transfer split in two, so that a field's loads are in flight
// while the block transforms the previous one: yf_issue loads this thread's
// Q elements of the tile into registers, yf_commit writes them to the tile.
template <int Q>
__device__ __forceinline__ void yf_issue(cplx (&r)[Q], const cplx* F, long P, long L0, int nv,
                                         int N) {
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = threadIdx.x + q * 64 * YF_LINES;
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    r[q] = (e < N * YF_LINES && c < nv) ? gld(F + (long)j * P + L0 + c) : mkc(0.0, 0.0);
  }
}
template <int Q>
__device__ __forceinline__ void yf_commit(cplx* T, const cplx (&r)[Q], int N) {
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = threadIdx.x + q * 64 * YF_LINES;
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    if (e < N * YF_LINES) T[TileCol{c}(j)] = r[q];
  }
  __syncthreads();
}

// The same transfers for the inter-stage fields (rhg, rhv, rdome in; v, dv,
// dome out), stored FP32 when io32 (the mixed precision mode).  The y
// kernels take the storage as a template argument IO (the bits of
// Geom::io32), so that the FP64 kernels carry no FP32 path.
// An FP32 element travels as its 8-byte pattern in the .x of the register
// until yf_commit_io widens it, so that the load stays in flight like an
// FP64 one (a conversion right after the load would wait for it).
__device__ __forceinline__ double yf_raw64(const cplx* F, long idx) {
  return *(const SNF_AS1 double*)(reinterpret_cast<const cplxf*>(F) + idx);
}
__device__ __forceinline__ cplx yf_widen(double raw) {
  const long long b = __double_as_longlong(raw);
  return mkc((double)__int_as_float((int)(b & 0xffffffffLL)), (double)__int_as_float((int)(b >> 32)));
}
template <int Q>
__device__ __forceinline__ void yf_issue_io(cplx (&r)[Q], const cplx* F, long P, long L0, int nv,
                                            int N, int io32) {
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = threadIdx.x + q * 64 * YF_LINES;
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    const long idx = (long)j * P + L0 + c;
    if (e < N * YF_LINES && c < nv)
      r[q] = io32 ? mkc(yf_raw64(F, idx), 0.0) : gld(F + idx);
    else
      r[q] = io32 ? mkc(__longlong_as_double(0LL), 0.0) : mkc(0.0, 0.0);
  }
}
template <int Q>
__device__ __forceinline__ void yf_commit_io(cplx* T, const cplx (&r)[Q], int N, int io32) {
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = threadIdx.x + q * 64 * YF_LINES;
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    if (e < N * YF_LINES) T[TileCol{c}(j)] = io32 ? yf_widen(r[q].x) : r[q];
  }
  __syncthreads();
}
__device__ __forceinline__ void yf_fill_io(cplx* T, const cplx* F, long P, long L0, int nv, int N,
                                           int io32) {
  __syncthreads();
  for (int e = threadIdx.x; e < N * YF_LINES; e += blockDim.x) {
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    T[TileCol{c}(j)] = (c < nv) ? ld_io(F, (long)j * P + L0 + c, io32) : mkc(0.0, 0.0);
  }
  __syncthreads();
}
__device__ __forceinline__ void yf_flush_io(const cplx* T, cplx* F, long P, long L0, int nv, int N,
                                            int io32) {
  __syncthreads();
  for (int e = threadIdx.x; e < N * YF_LINES; e += blockDim.x) {
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    if (c < nv) st_io(F, (long)j * P + L0 + c, T[TileCol{c}(j)], io32);
  }
  __syncthreads();
}

// Transfers of a block's chunk W (the RK accumulators vorwk, phiwk at the
// block's L0) in the element order of the tile transfers (e = j YF_LINES + c).
template <int Q>
__device__ __forceinline__ void yf_issue_blk(cplx (&r)[Q], const cplx* W, int N) {
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = threadIdx.x + q * 64 * YF_LINES;
    r[q] = e < N * YF_LINES ? gld(W + e) : mkc(0.0, 0.0);
  }
}
__device__ __forceinline__ void yf_flush_blk(const cplx* T, cplx* W, int N) {
  __syncthreads();
  for (int e = threadIdx.x; e < N * YF_LINES; e += blockDim.x) {
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    gst(W + e, T[TileCol{c}(j)]);
  }
  __syncthreads();
}

// the same for a chunk stored FP32 when io32 (cplxf element e at W + e)
template <int Q>
__device__ __forceinline__ void yf_issue_blk_io(cplx (&r)[Q], const cplx* W, int N, int io32) {
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = threadIdx.x + q * 64 * YF_LINES;
    if (e < N * YF_LINES)
      r[q] = io32 ? mkc(yf_raw64(W, e), 0.0) : gld(W + e);
    else
      r[q] = io32 ? mkc(__longlong_as_double(0LL), 0.0) : mkc(0.0, 0.0);
  }
}
__device__ __forceinline__ void yf_flush_blk_io(const cplx* T, cplx* W, int N, int io32) {
  __syncthreads();
  for (int e = threadIdx.x; e < N * YF_LINES; e += blockDim.x) {
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    st_io(W, e, T[TileCol{c}(j)], io32);
  }
  __syncthreads();
}
__device__ __forceinline__ void yf_fill_blk_io(cplx* T, const cplx* W, int N, int io32) {
  __syncthreads();
  for (int e = threadIdx.x; e < N * YF_LINES; e += blockDim.x) {
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    T[TileCol{c}(j)] = ld_io(W, e, io32);
  }
  __syncthreads();
}

__device__ __forceinline__ void yf_fill_blk(cplx* T, const cplx* W, int N) {
  __syncthreads();
  for (int e = threadIdx.x; e < N * YF_LINES; e += blockDim.x) {
    const int j = e / YF_LINES, c = e - j * YF_LINES;
    T[TileCol{c}(j)] = gld(W + e);
  }
  __syncthreads();
}

// A second field transfer split into issue and commit.  LIVE: as yf_issue /
// yf_commit, the loads in flight in registers in between.  Otherwise the
// issue only notes the source and the commit loads it (yf_fill): for plans
// of 9 rows per lane, whose registers do not hold a second field in flight
// beside the transforms.
template <int Q, bool LIVE> struct YPf;
template <int Q> struct YPf<Q, true> {
  cplx r[Q];
  __device__ __forceinline__ void issue(const cplx* F, long P, long L0, int nv, int N) {
    yf_issue(r, F, P, L0, nv, N);
  }
  __device__ __forceinline__ void issue_io(const cplx* F, long P, long L0, int nv, int N, int io32) {
    yf_issue_io(r, F, P, L0, nv, N, io32);
  }
  __device__ __forceinline__ void issue_blk(const cplx* W, int N) { yf_issue_blk(r, W, N); }
  __device__ __forceinline__ void issue_blk_io(const cplx* W, int N, int io32) {
    yf_issue_blk_io(r, W, N, io32);
  }
  __device__ __forceinline__ void commit(cplx* T, int N) { yf_commit(T, r, N); }
  __device__ __forceinline__ void commit_io(cplx* T, int N, int io32) { yf_commit_io(T, r, N, io32); }
};
template <int Q> struct YPf<Q, false> {
  const cplx* F;
  long P, L0;
  int nv;
  bool blk;
  __device__ __forceinline__ void issue(const cplx* F_, long P_, long L0_, int nv_, int) {
    F = F_;
    P = P_;
    L0 = L0_;
    nv = nv_;
    blk = false;
  }
  __device__ __forceinline__ void issue_io(const cplx* F_, long P_, long L0_, int nv_, int N, int) {
    issue(F_, P_, L0_, nv_, N);
  }
  __device__ __forceinline__ void issue_blk(const cplx* W, int) {
    F = W;
    blk = true;
  }
  __device__ __forceinline__ void issue_blk_io(const cplx* W, int N, int) { issue_blk(W, N); }
  __device__ __forceinline__ void commit(cplx* T, int N) {
    if (blk)
      yf_fill_blk(T, F, N);
    else
      yf_fill(T, F, P, L0, nv, N);
  }
  __device__ __forceinline__ void commit_io(cplx* T, int N, int io32) {
    if (blk)
      yf_fill_blk_io(T, F, N, io32);
    else
      yf_fill_io(T, F, P, L0, nv, N, io32);
  }
};
// whether a plan's kernels keep the second field transfer in flight
template <class PL> struct YPfLive {
  static constexpr bool value = PL::Q <= 8;
};

// exp(i w_m) of mode m: exp(2 pi i m/N) exp(i theta/N)
__device__ __forceinline__ cplx yf_mode(const cplx* rt, int m, cplx eth) {
  return cmul(cconj(gld(rt + m)), eth);
}

// Multipliers of the twisted d/dy, d2/dy2 and compact filter for every kx
// and mode m at the twist shp[kx]: out[(kx N + m) YS + {0,1,2}] = R1/B1,
// R2/B2, R_f/B_f at w_m = (2 pi m + arg shp)/N.  (The Helmholtz multiplier
// of a line is 1/(R2/B2 - k2).)
__device__ __forceinline__ void k_ysym_body(int N, YCoef c, const cplx* __restrict__ shp, const cplx* __restrict__ rt,
       cplx* __restrict__ out, const DevScalars* sc) {
  if (sc && sc->halted) return;
  const int kx = blockIdx.x, m = blockIdx.y * blockDim.x + threadIdx.x;
  if (m >= N) return;
  const cplx t = gld(shp + kx);
  const cplx e1 = yf_mode(rt, m, yf_expi(atan2(t.y, t.x) / (double)N));
  const C7 rh1 = c7_of(c.rh1), pr1 = c7_of(c.pr1), rh2 = c7_of(c.rh2), pr2 = c7_of(c.pr2);
  cplx* o = out + ((long)kx * N + m) * YS;
  gst(o, yf_div(yf_symbol(rh1, e1), yf_symbol(pr1, e1)));
  gst(o + 1, yf_div(yf_symbol(rh2, e1), yf_symbol(pr2, e1)));
  gst(o + 2, mkc(yf_filter_symbol(e1), 0.0));
}

template <class PL, int IO>
__device__ __forceinline__ void k_yf_explicit_body(YArgs a, int rk) {
  constexpr int Q = PL::Q;
  const int N = PL::direct ? a.g.my : PL::NC;
  if (a.sc->halted) return;
  extern __shared__ cplx T[];
  const Geom& g = a.g;
  const long P = SNF_CW;   // y stride of a line
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // blocks: YF_LINES consecutive kz lines of one kx (nch chunks per kx)
  const int nch = (g.mz + YF_LINES - 1) / YF_LINES;
  const long nbl = (long)g.kxn * nch;
  // ---- (kx, kz) lines; every wave takes part in every block transfer.
  // The blocks of one kx share its multiplier rows: keep them on one XCD.
  // A line block issues its vor loads first, so that they are in flight
  // while the block stages its tables.
  const long lb = xcd_block(blockIdx.x, nbl);
  const int i = (int)(lb / nch), ch = (int)(lb - (long)i * nch);
  const long L0 = sline(g, i, YF_LINES * ch);   // the block's chunk: my * YF_LINES elements
  const int nv = min(YF_LINES, g.mz - YF_LINES * ch);   // lines of this block
  cplx pv[Q];
  if (blockIdx.x != nbl) yf_issue(pv, a.vor, P, L0, nv, N);
  // LDS: tile [N][YF_LINES], one column per wave, the plan's table (the
  // N-th roots or twiddles), the multiplier rows of this block's kx at the
  // old and new twist
  cplx* rt = T + 2 * YF_LINES * N;
  for (int t = threadIdx.x; t < PL::tw_len(N); t += blockDim.x) rt[t] = gld((PL::direct ? a.twy : a.tpy) + t);
  __syncthreads();
  const int ir = rk - 1;
  double dt, f1, f2;
  substep_factors(a.sc, g.s, rk, dt, f1, f2);
  const double fac = 1.0 / g.Re;
  const double gam = c_gama[ir], zet = c_zeta[ir];
  const double invN = 1.0 / (double)N;
  const TileCol tc{w};
  const OwnCol ac{YF_LINES * N + w * N};
  cplx x[Q];

  if (blockIdx.x == nbl) {
    // ---- 00 lines: U (wave 0) and W (wave 1), real, untwisted (m=1 calls)
    if (w >= 2) return;
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
        T[ac(j)] = x[q];
      }
    }
    PL::template fft_wave<-1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) {
        T[ac(m)] = cmul(T[ac(m)], gld(sy + YS * m + 1));
      }
    }
    PL::template fft_wave<1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) {
        const double rfj = gldd(rf + j) + fac * (T[ac(j)].x * invN);
        double wkj = (rk == 1) ? gam * rfj : gld(uwk + j).x + gam * rfj;
        const double un = x[q].x + dt * wkj;
        if (rk != 3) wkj = rfj * zet;
        gst(u + j, mkc(un, 0.0));
        gst(uwk + j, mkc(wkj, 0.0));
        T[ac(j)] = mkc(un, 0.0);
      }
    }
    PL::template fft_wave<-1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) {
        T[ac(m)] = cmul(T[ac(m)], gld(sy + YS * m));
      }
    }
    PL::template fft_wave<1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) gstd(dd + j, T[ac(j)].x * invN);
    }
    return;
  }

  cplx* vwk_blk = a.vorwk + L0;    // the block's chunk of the accumulators
  cplx* pwk_blk = a.phiwk + L0;
  const int k = min(YF_LINES * ch + w, g.mz - 1);
  const double kx = g.alp * (double)i, kzv = kz_of(k, g.nz, g.mz, g.gam);
  const double k2 = kx * kx + kzv * kzv;
  const bool is00 = (i == 0 && k == 0);
  const cplx so = gld(a.shp_old + i), sn = gld(a.shp_new + i);
  const double tho = atan2(so.y, so.x) * invN, thn = atan2(sn.y, sn.x) * invN;
  // d/dy, d2/dy2 at the old and the new twist
  cplx* syo_s = rt + PL::tw_len(N);
  cplx* syn_s = syo_s + ys_lds<PL>(N);
  const auto syo = ys_row<PL>(syo_s, a.ysym_old, i, N);
  const auto syn = ys_row<PL>(syn_s, a.ysym_new, i, N);
  // phase tables: old twist, new twist, the two moving-frame maps
  cplx* pho = syn_s + ys_lds<PL>(N);
  cplx* phn = pho + YPH_N(Q);
  cplx* pm1 = phn + YPH_N(Q);
  cplx* pm2 = pm1 + YPH_N(Q);
  {
    const double pa[4] = {tho, thn, kx * f1 * g.dy, kx * f2 * g.dy};
    const double pa0[4] = {0.0, 0.0, -0.5 * kx * f1 * g.Ly, -0.5 * kx * f2 * g.Ly};
    for (int t = threadIdx.x; t < 4 * YPH_N(Q); t += blockDim.x) pho[t] = yph_entry(t, Q, pa, pa0);
  }

  // ===== omega_y chain =====
  // two fields' loads stay in flight (pfa, pfb; pfb where YPfLive) while the block transforms;
  // each is issued about two stages before the tile needs it
  cplx pfa[Q];
  YPf<Q, YPfLive<PL>::value> pfb;
  yf_commit(T, pv, N);                     // vor
  yf_issue_io(pfa, a.rhg, P, L0, nv, N, IO & 2);
  if (rk > 1) pfb.issue_blk(vwk_blk, N);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        x[q] = T[tc(j)];
        T[ac(j)] = cmul(x[q], e);
      }
    }
  }
  PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s2 = syo.d2(m);
      const cplx mul = is00 ? mkc(0.0, 0.0) : crmul(fac * invN, mkc(s2.x - k2, s2.y));
      T[ac(m)] = cmul(T[ac(m)], mul);
    }
  }
  PL::template fft<1>(T, ac, rt, N);
  yf_commit_io(T, pfa, N, IO & 2);  // rhg
  yf_issue_io(pfa, a.rdome, P, L0, nv, N, IO & 2);
  {
    YPh<false> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[ac(j)] = cadd(T[tc(j)], cmul(T[ac(j)], e));     // hg
    }
  }
  if (rk > 1) pfb.commit(T, N);       // vorwk
  pfb.issue(a.phi, P, L0, nv, N);
  {
    YPh<false> m1(pm1, lane);
    YPh<false> m2(pm2, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx p1 = m1.at(q), p2 = m2.at(q);
      if (j < N) {
        const cplx hgj = T[ac(j)];
        cplx vwk = (rk == 1) ? crmul(gam, hgj) : cadd(T[tc(j)], crmul(gam, hgj));
        const cplx vn = cadd(x[q], crmul(dt, vwk));
        if (rk != 3) vwk = cmul(crmul(zet, hgj), p1);
        x[q] = cmul(vn, p2);
        T[tc(j)] = vwk;
      }
    }
  }
  // the third substep's accumulators are never read (a step starts afresh)
  if (rk != 3) yf_flush_blk(T, vwk_blk, N);
  {
    YPh<true> ph(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        T[tc(j)] = x[q];
        T[ac(j)] = cmul(x[q], e);
      }
    }
  }
  yf_flush(T, a.vor, P, L0, nv, N);
  PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s1 = syn.d1(m);
      T[ac(m)] = cmul(T[ac(m)], crmul(invN, s1));
    }
  }
  PL::template fft<1>(T, ac, rt, N);
  {
    YPh<false> ph(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[tc(j)] = cmul(T[ac(j)], e);
    }
  }
  yf_flush_io(T, a.dome, P, L0, nv, N, IO & 1);

  // ===== phi chain =====
  yf_commit_io(T, pfa, N, IO & 2);  // rdome
  yf_issue_io(pfa, a.rhv, P, L0, nv, N, IO & 2);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[ac(j)] = cmul(T[tc(j)], e);
    }
  }
  PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s1 = syo.d1(m);
      T[ac(m)] = cmul(T[ac(m)], crmul(-invN, s1));
    }
  }
  pfb.commit(T, N);                    // phi
  if (rk > 1) pfb.issue_blk(pwk_blk, N);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        x[q] = T[tc(j)];
        T[tc(j)] = cmul(x[q], e);
      }
    }
  }
  PL::template fft<-1>(T, tc, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s2 = syo.d2(m);
      const cplx mul = is00 ? mkc(0.0, 0.0) : crmul(fac * invN, mkc(s2.x - k2, s2.y));
      T[ac(m)] = cadd(T[ac(m)], cmul(T[tc(m)], mul));
    }
  }
  PL::template fft<1>(T, ac, rt, N);
  yf_commit_io(T, pfa, N, IO & 2);  // rhv
  {
    YPh<false> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[ac(j)] = cadd(T[tc(j)], cmul(T[ac(j)], e));     // hv
    }
  }
  if (rk > 1) pfb.commit(T, N);       // phiwk
  {
    YPh<false> m1(pm1, lane);
    YPh<false> m2(pm2, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx p1 = m1.at(q), p2 = m2.at(q);
      if (j < N) {
        const cplx hvj = T[ac(j)];
        cplx pwk = (rk == 1) ? crmul(gam, hvj) : cadd(T[tc(j)], crmul(gam, hvj));
        const cplx pn = cadd(x[q], crmul(dt, pwk));
        if (rk != 3) pwk = cmul(crmul(zet, hvj), p1);
        x[q] = cmul(pn, p2);
        T[tc(j)] = pwk;
      }
    }
  }
  if (rk != 3) yf_flush_blk(T, pwk_blk, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = lane + 64 * q;
    if (j < N) T[tc(j)] = x[q];
  }
  yf_flush(T, a.phi, P, L0, nv, N);
  // v = lapcdy(phi) and dv = D1(v) with the new twist; the (0,0) mode keeps
  // its v (lapcdy is not called for it) and only takes dv = D1(v)
  if (L0 == 0) yf_fill_io(T, a.v, P, L0, nv, N, IO & 1);   // block of the (0,0) mode
  // every wave of the block reaches the same transforms in the same order
  // (the matrix-core plan transforms the block's columns together); the
  // (0,0) mode takes no Helmholtz solve and leaves out the last transform
  {
    YPh<true> ph(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        if (is00) x[q] = T[tc(j)];             // v of the (0,0) mode, unchanged
        T[ac(j)] = cmul(x[q], e);
      }
    }
  }
  PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s1 = syn.d1(m), s2 = syn.d2(m);
      if (!is00) {
        const cplx H = yf_div(mkc(invN, 0.0), mkc(s2.x - k2, s2.y));   // lapcdy
        const cplx vh = cmul(T[ac(m)], H);
        T[tc(m)] = cmul(vh, s1);
        T[ac(m)] = vh;
      } else {
        T[ac(m)] = cmul(T[ac(m)], crmul(invN, s1));
      }
    }
  }
  PL::template fft<1>(T, ac, rt, N);
  PL::template fft<1>(T, tc, rt, N, !is00);
  {
    YPh<false> ph(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        if (!is00) {
          x[q] = cmul(T[ac(j)], e);            // v
          T[tc(j)] = cmul(T[tc(j)], e);        // dv
        } else {
          T[tc(j)] = cmul(T[ac(j)], e);        // dv
        }
      }
    }
  }
  yf_flush_io(T, a.dv, P, L0, nv, N, IO & 1);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = lane + 64 * q;
    if (j < N) T[tc(j)] = x[q];
  }
  yf_flush_io(T, a.v, P, L0, nv, N, IO & 1);
}
template <class PL, int IO>
__global__ void __launch_bounds__(64 * YF_LINES) __attribute__((amdgpu_waves_per_eu(2)))
k_yf_explicit(YArgs a, int rk) {
  k_yf_explicit_body<PL, IO>(a, rk);
}

// Implicit RK3 substep of the y stage (cross.f90:797-990) in the same block
// layout as k_yf_explicit, with the Helmholtz solves (lapcdy with k2 + dtr1,
// diffy_shear_hepta.f90:466-525) as twisted-Fourier multipliers
// 1 / (R2/B2 - k2 - dtr1).  Per line, old twist theta, new twist theta':
//   omega_y: vwk = facb (gama hg [+ vorwk] + alpha/Re (D2 - k2) vor + vor/dt)
//            mapped by m2; vor = Helm'(vwk); dome = D1'(vor);
//            vorwk = zeta hg m1 (rk 1, 2)
//   phi:     hv = rhv - D1(rdome); pwk as vwk with hv, phi;
//            phi = Helm'(pwk); v = lapcdy'(phi); dv = D1'(v);
//            phiwk = zeta hv m1 (rk 1, 2)
// The (0,0) line keeps vor, phi and v (cross.f90:866, 952) and takes dome,
// dv = D1 of them.  U, W (m=1 calls, last block) follow cross.f90:829-843,
// 858-863, 900-918.  13 line transforms; every field read and written once.
template <class PL, int IO>
__device__ __forceinline__ void k_yf_implicit_body(YArgs a, int rk) {
  constexpr int Q = PL::Q;
  const int N = PL::direct ? a.g.my : PL::NC;
  if (a.sc->halted) return;
  extern __shared__ cplx T[];
  const Geom& g = a.g;
  const long P = SNF_CW;   // y stride of a line
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nch = (g.mz + YF_LINES - 1) / YF_LINES;
  const long nbl = (long)g.kxn * nch;
  // a line block's vor loads are in flight while it stages its tables
  const long lb = xcd_block(blockIdx.x, nbl);
  const int i = (int)(lb / nch), ch = (int)(lb - (long)i * nch);
  const long L0 = sline(g, i, YF_LINES * ch);   // the block's chunk: my * YF_LINES elements
  const int nv = min(YF_LINES, g.mz - YF_LINES * ch);
  cplx pv[Q];
  if (blockIdx.x != nbl) yf_issue(pv, a.vor, P, L0, nv, N);
  cplx* rt = T + 2 * YF_LINES * N;
  for (int t = threadIdx.x; t < PL::tw_len(N); t += blockDim.x) rt[t] = gld((PL::direct ? a.twy : a.tpy) + t);
  __syncthreads();
  const int ir = rk - 1;
  double dt, f1, f2;
  substep_factors(a.sc, g.s, rk, dt, f1, f2);
  const double fva = c_alpha[ir] * (1.0 / g.Re);   // alpha(rk)/Re
  const double facb = -g.Re * c_ibeta[ir];
  const double rdt = 1.0 / dt;
  const double dtr1 = a.sc->dtr * c_ibeta[ir];
  const double gam = c_gama[ir], zet = c_zeta[ir];
  const double invN = 1.0 / (double)N;
  const TileCol tc{w};
  const OwnCol ac{YF_LINES * N + w * N};
  cplx x[Q];

  if (blockIdx.x == nbl) {
    // ---- 00 lines: U (wave 0) and W (wave 1), real, untwisted
    if (w >= 2) return;
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
        T[ac(j)] = x[q];
      }
    }
    PL::template fft_wave<-1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) T[ac(m)] = cmul(T[ac(m)], gld(sy + YS * m + 1));
    }
    PL::template fft_wave<1>(T, ac, rt, N);
    // wk = facb (gama rf [+ wk] + alpha/Re D2 U + U/dt)
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) {
        const double rfj = gldd(rf + j);
        double wkj = (rk == 1) ? gam * rfj : gld(uwk + j).x + gam * rfj;
        wkj = wkj + fva * (T[ac(j)].x * invN);
        wkj = facb * (wkj + rdt * x[q].x);
        T[ac(j)] = mkc(wkj, 0.0);
        if (rk != 3) gst(uwk + j, mkc(zet * rfj, 0.0));
      }
    }
    // U = lapcdy(wk, dtr1); dU/dy
    PL::template fft_wave<-1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) {
        const cplx s1 = gld(sy + YS * m), s2 = gld(sy + YS * m + 1);
        const cplx uh = cmul(T[ac(m)], yf_div(mkc(invN, 0.0), mkc(s2.x - dtr1, s2.y)));
        T[ac(m)] = uh;
        T[tc(m)] = cmul(uh, s1);
      }
    }
    PL::template fft_wave<1>(T, ac, rt, N);
    PL::template fft_wave<1>(T, tc, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) {
        gst(u + j, mkc(T[ac(j)].x, 0.0));
        gstd(dd + j, T[tc(j)].x);
      }
    }
    return;
  }

  cplx* vwk_blk = a.vorwk + L0;    // the block's chunk of the accumulators
  cplx* pwk_blk = a.phiwk + L0;
  const int k = min(YF_LINES * ch + w, g.mz - 1);
  const double kx = g.alp * (double)i, kzv = kz_of(k, g.nz, g.mz, g.gam);
  const double k2 = kx * kx + kzv * kzv;
  const bool is00 = (i == 0 && k == 0);
  const cplx so = gld(a.shp_old + i), sn = gld(a.shp_new + i);
  const double tho = atan2(so.y, so.x) * invN, thn = atan2(sn.y, sn.x) * invN;
  // d/dy, d2/dy2 at the old and the new twist
  cplx* syo_s = rt + PL::tw_len(N);
  cplx* syn_s = syo_s + ys_lds<PL>(N);
  const auto syo = ys_row<PL>(syo_s, a.ysym_old, i, N);
  const auto syn = ys_row<PL>(syn_s, a.ysym_new, i, N);
  cplx* pho = syn_s + ys_lds<PL>(N);
  cplx* phn = pho + YPH_N(Q);
  cplx* pm1 = phn + YPH_N(Q);
  cplx* pm2 = pm1 + YPH_N(Q);
  {
    const double pa[4] = {tho, thn, kx * f1 * g.dy, kx * f2 * g.dy};
    const double pa0[4] = {0.0, 0.0, -0.5 * kx * f1 * g.Ly, -0.5 * kx * f2 * g.Ly};
    for (int t = threadIdx.x; t < 4 * YPH_N(Q); t += blockDim.x) pho[t] = yph_entry(t, Q, pa, pa0);
  }
  // Helmholtz multiplier 1/(R2/B2 - k2 - dtr1) / N at the new twist
  auto helm = [&](int m) {
    const cplx s2 = syn.d2(m);
    return yf_div(mkc(invN, 0.0), mkc(s2.x - (k2 + dtr1), s2.y));
  };

  // ===== omega_y chain =====
  cplx pfa[Q];
  YPf<Q, YPfLive<PL>::value> pfb;
  yf_commit(T, pv, N);                     // vor
  yf_issue_io(pfa, a.rhg, P, L0, nv, N, IO & 2);
  if (rk > 1) pfb.issue_blk(vwk_blk, N);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        x[q] = T[tc(j)];
        T[ac(j)] = cmul(x[q], e);
      }
    }
  }
  PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s2 = syo.d2(m);
      const cplx mul = is00 ? mkc(0.0, 0.0) : crmul(fva * invN, mkc(s2.x - k2, s2.y));
      T[ac(m)] = cmul(T[ac(m)], mul);
    }
  }
  PL::template fft<1>(T, ac, rt, N);
  // own: alpha/Re (D2 - k2) vor + vor/dt [+ vorwk]
  if (rk > 1) pfb.commit(T, N);       // vorwk
  {
    YPh<false> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        cplx o = cadd(cmul(T[ac(j)], e), crmul(rdt, x[q]));
        if (rk > 1) o = cadd(T[tc(j)], o);
        T[ac(j)] = o;
      }
    }
  }
  yf_commit_io(T, pfa, N, IO & 2);  // rhg
  yf_issue_io(pfa, a.rdome, P, L0, nv, N, IO & 2);
  pfb.issue(a.phi, P, L0, nv, N);
  {
    YPh<false> m1(pm1, lane);
    YPh<false> m2(pm2, lane);
    YPh<true> en(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx p1 = m1.at(q), p2 = m2.at(q), e = en.at(q);
      if (j < N) {
        const cplx hgj = T[tc(j)];
        const cplx vwk = cmul(crmul(facb, cadd(T[ac(j)], crmul(gam, hgj))), p2);
        T[ac(j)] = cmul(is00 ? x[q] : vwk, e);      // to the Helmholtz solve (00: D1 of vor)
        if (rk != 3) T[tc(j)] = cmul(crmul(zet, hgj), p1);
      }
    }
  }
  if (rk != 3) yf_flush_blk(T, vwk_blk, N);
  PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx vh = is00 ? crmul(invN, T[ac(m)]) : cmul(T[ac(m)], helm(m));
      T[ac(m)] = vh;
      T[tc(m)] = cmul(vh, syn.d1(m));
    }
  }
  PL::template fft<1>(T, ac, rt, N);
  PL::template fft<1>(T, tc, rt, N);
  {
    YPh<false> en(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = en.at(q);
      if (j < N) {
        if (!is00) x[q] = cmul(T[ac(j)], e);   // vor (the (0,0) line keeps its own)
        T[tc(j)] = cmul(T[tc(j)], e);          // dome
      }
    }
  }
  yf_flush_io(T, a.dome, P, L0, nv, N, IO & 1);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = lane + 64 * q;
    if (j < N) T[tc(j)] = x[q];
  }
  yf_flush(T, a.vor, P, L0, nv, N);

  // ===== phi chain =====
  pfb.commit(T, N);                    // phi
  pfb.issue_io(a.rhv, P, L0, nv, N, IO & 2);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        x[q] = T[tc(j)];
        T[ac(j)] = cmul(x[q], e);
      }
    }
  }
  PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s2 = syo.d2(m);
      const cplx mul = is00 ? mkc(0.0, 0.0) : crmul(fva * invN, mkc(s2.x - k2, s2.y));
      T[ac(m)] = cmul(T[ac(m)], mul);
    }
  }
  PL::template fft<1>(T, ac, rt, N);
  {
    YPh<false> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[ac(j)] = cadd(cmul(T[ac(j)], e), crmul(rdt, x[q]));
    }
  }
  // the (0,0) line keeps phi: x holds it until the flush
  yf_commit_io(T, pfa, N, IO & 2);  // rdome
  if (rk > 1) yf_issue_blk(pfa, pwk_blk, N);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[tc(j)] = cmul(T[tc(j)], e);
    }
  }
  PL::template fft<-1>(T, tc, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) T[tc(m)] = cmul(T[tc(m)], crmul(invN, syo.d1(m)));
  }
  PL::template fft<1>(T, tc, rt, N);
  // pfb holds rhv; D1(rdome) goes to registers d, the tile turns over
  cplx d[Q];
  {
    YPh<false> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) d[q] = cmul(T[tc(j)], e);
    }
  }
  if (rk > 1) {
    yf_commit(T, pfa, N);                  // phiwk
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) T[ac(j)] = cadd(T[tc(j)], T[ac(j)]);
    }
  }
  pfb.commit_io(T, N, IO & 2);  // rhv
  {
    YPh<false> m1(pm1, lane);
    YPh<false> m2(pm2, lane);
    YPh<true> en(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx p1 = m1.at(q), p2 = m2.at(q), e = en.at(q);
      if (j < N) {
        const cplx hvj = csub(T[tc(j)], d[q]);                 // hv = rhv - D1(rdome)
        const cplx pwk = cmul(crmul(facb, cadd(T[ac(j)], crmul(gam, hvj))), p2);
        T[ac(j)] = cmul(pwk, e);
        if (rk != 3) T[tc(j)] = cmul(crmul(zet, hvj), p1);
      }
    }
  }
  if (rk != 3) yf_flush_blk(T, pwk_blk, N);
  // phi = Helm'(pwk); v = lapcdy'(phi), dv = D1'(v)
  PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx ph = cmul(T[ac(m)], helm(m));
      T[ac(m)] = ph;
      T[tc(m)] = ph;
    }
  }
  PL::template fft<1>(T, tc, rt, N);
  {
    YPh<false> en(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = en.at(q);
      if (j < N && !is00) x[q] = cmul(T[tc(j)], e);   // phi
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = lane + 64 * q;
    if (j < N) T[tc(j)] = x[q];
  }
  yf_flush(T, a.phi, P, L0, nv, N);
  if (L0 == 0) yf_fill_io(T, a.v, P, L0, nv, N, IO & 1);   // block of the (0,0) mode
  if (!is00) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) {
        const cplx s1 = syn.d1(m), s2 = syn.d2(m);
        // own holds Helm'(pwk) / N with the 1/N already applied once: the
        // lapcdy multiplier B2/(R2 - k2 B2) without a second 1/N
        const cplx vh = yf_div(T[ac(m)], mkc(s2.x - k2, s2.y));
        T[tc(m)] = cmul(vh, s1);
        T[ac(m)] = vh;
      }
    }
  } else {
    YPh<true> en(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = en.at(q);
      if (j < N) {
        x[q] = T[tc(j)];                     // v of the (0,0) mode, unchanged
        T[ac(j)] = cmul(x[q], e);
      }
    }
  }
  // every wave reaches this transform (the matrix-core plan transforms the
  // block's columns together); only the (0,0) mode's column takes part
  PL::template fft<-1>(T, ac, rt, N, is00);
  if (is00) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) {
        T[tc(m)] = cmul(T[ac(m)], crmul(invN, syn.d1(m)));
        T[ac(m)] = crmul(invN, T[ac(m)]);
      }
    }
  }
  PL::template fft<1>(T, ac, rt, N);
  PL::template fft<1>(T, tc, rt, N);
  {
    YPh<false> en(phn, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = en.at(q);
      if (j < N) {
        if (!is00) x[q] = cmul(T[ac(j)], e);   // v
        T[tc(j)] = cmul(T[tc(j)], e);          // dv
      }
    }
  }
  yf_flush_io(T, a.dv, P, L0, nv, N, IO & 1);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = lane + 64 * q;
    if (j < N) T[tc(j)] = x[q];
  }
  yf_flush_io(T, a.v, P, L0, nv, N, IO & 1);
}
template <class PL, int IO>
__global__ void __launch_bounds__(64 * YF_LINES) __attribute__((amdgpu_waves_per_eu(2)))
k_yf_implicit(YArgs a, int rk) {
  k_yf_implicit_body<PL, IO>(a, rk);
}

// Start of a cross call on the spectral y stage (cross.f90:408-445), the
// role of k_y_first: at the current twist, v = lapcdy(phi) (the (0,0) line
// keeps its v), dv = D1(v), dome = D1(vor); dU/dy, dW/dy (cross.f90:1198-1205)
// in the last block.  Same block layout as k_yf_explicit; 6 line transforms.
template <class PL, int IO>
__global__ void __launch_bounds__(64 * YF_LINES) __attribute__((amdgpu_waves_per_eu(2)))
k_yf_first(YArgs a) {
  constexpr int Q = PL::Q;
  const int N = PL::direct ? a.g.my : PL::NC;
  if (a.sc->halted) return;
  extern __shared__ cplx T[];
  const Geom& g = a.g;
  const long P = SNF_CW;   // y stride of a line
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nch = (g.mz + YF_LINES - 1) / YF_LINES;
  const long nbl = (long)g.kxn * nch;
  cplx* rt = T + 2 * YF_LINES * N;
  for (int t = threadIdx.x; t < PL::tw_len(N); t += blockDim.x) rt[t] = gld((PL::direct ? a.twy : a.tpy) + t);
  __syncthreads();
  const double invN = 1.0 / (double)N;
  const TileCol tc{w};
  const OwnCol ac{YF_LINES * N + w * N};
  cplx x[Q];

  if (blockIdx.x == nbl) {
    // ---- dU/dy (wave 0), dW/dy (wave 1): real, untwisted
    if (w >= 2) return;
    const cplx* u = w == 0 ? a.u00c : a.w00c;
    double* dd = w == 0 ? a.dudy : a.dwdy;
    const cplx* sy = a.ysym_old;   // kx = 0 row: untwisted
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) T[ac(j)] = gld(u + j);
    }
    PL::template fft_wave<-1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) T[ac(m)] = cmul(T[ac(m)], crmul(invN, gld(sy + YS * m)));
    }
    PL::template fft_wave<1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) gstd(dd + j, T[ac(j)].x);
    }
    return;
  }

  const long lb = xcd_block(blockIdx.x, nbl);
  const int i = (int)(lb / nch), ch = (int)(lb - (long)i * nch);
  const long L0 = sline(g, i, YF_LINES * ch);   // the block's chunk: my * YF_LINES elements
  const int nv = min(YF_LINES, g.mz - YF_LINES * ch);
  const int k = min(YF_LINES * ch + w, g.mz - 1);
  const double kx = g.alp * (double)i, kzv = kz_of(k, g.nz, g.mz, g.gam);
  const double k2 = kx * kx + kzv * kzv;
  const bool is00 = (i == 0 && k == 0);
  const cplx so = gld(a.shp_old + i);
  const double tho = atan2(so.y, so.x) * invN;
  cplx* sy_s = rt + PL::tw_len(N);         // d/dy, d2/dy2 at the current twist
  const auto sy = ys_row<PL>(sy_s, a.ysym_old, i, N);
  cplx* pho = sy_s + ys_lds<PL>(N);
  {
    const double pa[1] = {tho}, pa0[1] = {0.0};
    for (int t = threadIdx.x; t < YPH_N(Q); t += blockDim.x) pho[t] = yph_entry(t, Q, pa, pa0);
  }

  // dome = D1(vor)
  yf_fill(T, a.vor, P, L0, nv, N);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[ac(j)] = cmul(T[tc(j)], e);
    }
  }
  PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) T[ac(m)] = cmul(T[ac(m)], crmul(invN, sy.d1(m)));
  }
  PL::template fft<1>(T, ac, rt, N);
  {
    YPh<false> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[tc(j)] = cmul(T[ac(j)], e);
    }
  }
  yf_flush_io(T, a.dome, P, L0, nv, N, IO & 1);

  // v = lapcdy(phi) (the (0,0) line keeps v), dv = D1(v)
  if (L0 == 0) {                 // block of the (0,0) mode
    yf_fill_io(T, a.v, P, L0, nv, N, IO & 1);
    if (is00) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int j = lane + 64 * q;
        if (j < N) x[q] = T[tc(j)];
      }
    }
  }
  yf_fill(T, a.phi, P, L0, nv, N);
  {
    YPh<true> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) T[ac(j)] = cmul(is00 ? x[q] : T[tc(j)], e);
    }
  }
  PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = lane + 64 * q;
    if (m < N) {
      const cplx s1 = sy.d1(m), s2 = sy.d2(m);
      const cplx vh = is00 ? crmul(invN, T[ac(m)])
                           : cmul(T[ac(m)], yf_div(mkc(invN, 0.0), mkc(s2.x - k2, s2.y)));
      T[tc(m)] = cmul(vh, s1);
      T[ac(m)] = vh;
    }
  }
  PL::template fft<1>(T, ac, rt, N);
  PL::template fft<1>(T, tc, rt, N);
  {
    YPh<false> ph(pho, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      const cplx e = ph.at(q);
      if (j < N) {
        if (!is00) x[q] = cmul(T[ac(j)], e);   // v
        T[tc(j)] = cmul(T[tc(j)], e);          // dv
      }
    }
  }
  yf_flush_io(T, a.dv, P, L0, nv, N, IO & 1);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = lane + 64 * q;
    if (j < N) T[tc(j)] = x[q];
  }
  yf_flush_io(T, a.v, P, L0, nv, N, IO & 1);
}

// iydealiasing=3: the compact filter (filterc, filter_shear_hepta.f90:128-165)
// applied in place to vor and phi at the current twist and to U, W
// (cross.f90:214-225, 476-488), then dU/dy, dW/dy of the filtered means
// (read by the x-z stage, cross.f90:1198-1205).  Same block layout as
// k_yf_explicit.
template <class PL>
__global__ void __launch_bounds__(64 * YF_LINES) __attribute__((amdgpu_waves_per_eu(2)))
k_yf_filter(YArgs a) {
  constexpr int Q = PL::Q;
  const int N = PL::direct ? a.g.my : PL::NC;
  if (a.sc->halted) return;
  extern __shared__ cplx T[];
  const Geom& g = a.g;
  const long P = SNF_CW;   // y stride of a line
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nch = (g.mz + YF_LINES - 1) / YF_LINES;
  const long nbl = (long)g.kxn * nch;
  cplx* rt = T + 2 * YF_LINES * N;
  for (int t = threadIdx.x; t < PL::tw_len(N); t += blockDim.x) rt[t] = gld((PL::direct ? a.twy : a.tpy) + t);
  __syncthreads();
  const double invN = 1.0 / (double)N;
  const TileCol tc{w};
  const OwnCol ac{YF_LINES * N + w * N};
  if (blockIdx.x == nbl) {
    if (w >= 2) return;
    cplx* u = w == 0 ? a.u00c : a.w00c;
    double* dd = w == 0 ? a.dudy : a.dwdy;
    const cplx* sy = a.ysym_old;   // kx = 0 row: untwisted
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) T[ac(j)] = gld(u + j);
    }
    PL::template fft_wave<-1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) T[ac(m)] = crmul(invN * gld(sy + YS * m + 2).x, T[ac(m)]);
    }
    PL::template fft_wave<1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) {
        const cplx uf = mkc(T[ac(j)].x, 0.0);
        gst(u + j, uf);
        T[ac(j)] = uf;
      }
    }
    PL::template fft_wave<-1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) T[ac(m)] = cmul(T[ac(m)], crmul(invN, gld(sy + YS * m)));
    }
    PL::template fft_wave<1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = lane + 64 * q;
      if (j < N) gstd(dd + j, T[ac(j)].x);
    }
    return;
  }
  const long lb = xcd_block(blockIdx.x, nbl);
  const int i = (int)(lb / nch), ch = (int)(lb - (long)i * nch);
  const long L0 = sline(g, i, YF_LINES * ch);   // the block's chunk: my * YF_LINES elements
  const int nv = min(YF_LINES, g.mz - YF_LINES * ch);
  const cplx so = gld(a.shp_old + i);
  const double tho = atan2(so.y, so.x) * invN;
  const cplx* sy = a.ysym_old + (long)i * N * YS;
  for (int f = 0; f < 2; ++f) {
    cplx* F = f == 0 ? a.vor : a.phi;
    yf_fill(T, F, P, L0, nv, N);
    {
      YPhase ph(-tho, 0.0, lane);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int j = lane + 64 * q;
        const cplx e = ph.next();
        if (j < N) T[ac(j)] = cmul(T[tc(j)], e);
      }
    }
    PL::template fft<-1>(T, ac, rt, N);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int m = lane + 64 * q;
      if (m < N) T[ac(m)] = crmul(invN * gld(sy + YS * m + 2).x, T[ac(m)]);
    }
    PL::template fft<1>(T, ac, rt, N);
    {
      YPhase ph(tho, 0.0, lane);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int j = lane + 64 * q;
        const cplx e = ph.next();
        if (j < N) T[tc(j)] = cmul(T[ac(j)], e);
      }
    }
    yf_flush(T, F, P, L0, nv, N);
  }
}

// box-table forms (batched ensembles): block z takes the arguments of box z
template <class PL, int IO>
__global__ void __launch_bounds__(64 * YF_LINES) __attribute__((amdgpu_waves_per_eu(2)))
k_yf_explicit_tab(const YArgs* __restrict__ t, int rk) {
  k_yf_explicit_body<PL, IO>(t[blockIdx.z], rk);
}
template <class PL, int IO>
__global__ void __launch_bounds__(64 * YF_LINES) __attribute__((amdgpu_waves_per_eu(2)))
k_yf_implicit_tab(const YArgs* __restrict__ t, int rk) {
  k_yf_implicit_body<PL, IO>(t[blockIdx.z], rk);
}
