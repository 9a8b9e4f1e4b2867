// Spectral application of the compact y-operators on twisted-periodic lines.
//
// The compact schemes of diffy_shear_hepta.f90 are constant-coefficient
// stencils B x = R u on a line whose extension is u(j + N) = shp u(j)
// (derivyc 358-383, lapcdy 487-514).  Such a twisted-cyclic operator is
// diagonal in the twisted Fourier modes u_j = exp(i w j), w = (theta + 2 pi m)/N,
// theta = arg shp: it multiplies mode m by
//     d/dy   : R1(w) / B1(w)                  (rh1, pr1)
//     d2/dy2 : R2(w) / B2(w)                  (rh2, pr2)
//     lapcdy : B2(w) / (R2(w) - k2 B2(w))     ((rh2 - k2 pr2) p = pr2 u)
// with C(w) = sum_p c(p) exp(i p w), p = -3..3.  The result equals the
// reference's cyclic L D L^H solve (modChol7cyclic_cmplx 578-755,
// solve7cyclic 900-954) up to rounding.  So a line is transformed once
// (x -> X_m = sum_j x_j exp(-i w_m j)), multiplied, and transformed back.
//
// Transforms: one wave per line, mixed-radix Stockham passes (radices 2, 3,
// 4, 5, 7, 8, 9, 11, 13, 19) in place on the line's column of an LDS tile,
// every pass a read-all / write-all of the wave's own column (no block
// barrier).
// Twiddles from a pass-major table staged in LDS: pass s > 0 (stride Ns,
// radix R) holds exp(-2 pi i r k / (Ns R)) at [r - 1][k], r < R, k < Ns, so
// that the lanes of a pass read consecutive entries (no bank conflicts);
// the first pass needs none.  yf_twiddle_table builds it on the host.
#pragma once
#include "snf_internal.h"
#include "snf_wfft.h"

#ifndef YF_LINES
#define YF_LINES 4        // lines (waves) per block; tile rows of 4 lines = 64 B
#endif

// Compile-time transform plan: the radices of the passes, in order.  A plan
// fixes N = prod(Rs) and the number of rows per lane Q = ceil(N / 64).
template <int... Rs> struct YPlanT;
template <> struct YPlanT<> {
  static constexpr int N = 1;
};
template <int R, int... Rs> struct YPlanT<R, Rs...> {
  static constexpr int N = R * YPlanT<Rs...>::N;
};

// Entries of the pass-major twiddle table of the plan R[0..np-1]:
// sum over the passes s > 0 of (R[s] - 1) * (R[0] ... R[s-1]); never above N.
inline int yf_twiddle_count(const int* R, int np) {
  int n = 0, Ns = R[0];
  for (int s = 1; s < np; ++s) {
    n += (R[s] - 1) * Ns;
    Ns *= R[s];
  }
  return n;
}

// Column accessors.  TileCol: line c of the block's transfer tile, rows of
// YF_LINES lines with the line index XORed with a function of j so that a
// wave walking one line down the rows touches distinct LDS banks: 16-byte
// writes are served 8 lanes (128 B) a cycle, reads 16 lanes (256 B), and the
// swizzle keeps both groups on distinct 16-byte slots.  OwnCol: a wave's
// private contiguous column.
__device__ __forceinline__ int yf_swz(int j) {
  if (YF_LINES == 2) return ((j >> 2) ^ (j >> 3)) & 1;
  if (YF_LINES == 4) return ((j >> 1) ^ (j >> 3)) & 3;
  return (j ^ (j >> 3)) & (YF_LINES - 1);
}
struct TileCol {
  int c;
  __device__ __forceinline__ int operator()(int j) const {
    return j * YF_LINES + (c ^ yf_swz(j));
  }
  // the same kind of column of line L of the block (this column is line w's)
  __device__ __forceinline__ TileCol peer(int L, int, int) const { return TileCol{L}; }
};
struct OwnCol {
  int base;
  __device__ __forceinline__ int operator()(int j) const { return base + j; }
  __device__ __forceinline__ OwnCol peer(int L, int w, int N) const { return OwnCol{base + (L - w) * N}; }
};

// ---- small odd DFTs by symmetric pairs: y_k = sum_j x_j exp(SIGN 2 pi i jk/R)
template <int R> struct OddRoots;
template <> struct OddRoots<5> {
  static constexpr double c[5] = {1.0, 0.3090169943749474241023, -0.8090169943749474241023,
                                  -0.8090169943749474241023, 0.3090169943749474241023};
  static constexpr double s[5] = {0.0, 0.9510565162951535721164, 0.5877852522924731291687,
                                  -0.5877852522924731291687, -0.9510565162951535721164};
};
template <> struct OddRoots<7> {
  static constexpr double c[7] = {1.0, 0.6234898018587335305250, -0.2225209339563144042889,
                                  -0.9009688679024191262361, -0.9009688679024191262361,
                                  -0.2225209339563144042889, 0.6234898018587335305250};
  static constexpr double s[7] = {0.0, 0.7818314824680298087084, 0.9749279121818236070181,
                                  0.4338837391175581204758, -0.4338837391175581204758,
                                  -0.9749279121818236070181, -0.7818314824680298087084};
};
template <> struct OddRoots<9> {
  static constexpr double c[9] = {1.0, 0.7660444431189780352024, 0.1736481776669303488517, -0.5,
                                  -0.9396926207859083840541, -0.9396926207859083840541, -0.5,
                                  0.1736481776669303488517, 0.7660444431189780352024};
  static constexpr double s[9] = {0.0, 0.6427876096865393263226, 0.9848077530122080593667,
                                  0.8660254037844386467637, 0.3420201433256687330441,
                                  -0.3420201433256687330441, -0.8660254037844386467637,
                                  -0.9848077530122080593667, -0.6427876096865393263226};
};
template <> struct OddRoots<11> {
  static constexpr double c[11] = {1.0, 0.8412535328311811688618, 0.4154150130018864255293,
                                   -0.1423148382732851404438, -0.6548607339452850640569,
                                   -0.9594929736144973898904, -0.9594929736144973898904,
                                   -0.6548607339452850640569, -0.1423148382732851404438,
                                   0.4154150130018864255293, 0.8412535328311811688618};
  static constexpr double s[11] = {0.0, 0.5406408174555975821076, 0.9096319953545183714117,
                                   0.9898214418809327323761, 0.7557495743542582837740,
                                   0.2817325568414296977114, -0.2817325568414296977114,
                                   -0.7557495743542582837740, -0.9898214418809327323761,
                                   -0.9096319953545183714117, -0.5406408174555975821076};
};

template <> struct OddRoots<13> {
  static constexpr double c[13] = {1.0, 0.8854560256532098959004, 0.5680647467311558025118,
                                   0.1205366802553230533491, -0.3546048870425356259696,
                                   -0.7485107481711010986346, -0.9709418174260520271570,
                                   -0.9709418174260520271570, -0.7485107481711010986346,
                                   -0.3546048870425356259696, 0.1205366802553230533491,
                                   0.5680647467311558025118, 0.8854560256532098959004};
  static constexpr double s[13] = {0.0, 0.4647231720437685456560, 0.8229838658936563945796,
                                   0.9927088740980539928008, 0.9350162426854148234398,
                                   0.6631226582407952023768, 0.2393156642875577671488,
                                   -0.2393156642875577671488, -0.6631226582407952023768,
                                   -0.9350162426854148234398, -0.9927088740980539928008,
                                   -0.8229838658936563945796, -0.4647231720437685456560};
};

template <> struct OddRoots<19> {
  static constexpr double c[19] = {
      1.0, 0.9458172417006346790197, 0.7891405093963935992190, 0.5469481581224268747118,
      0.2454854871407991489223, -0.08257934547233232460034, -0.4016954246529694575168,
      -0.6772815716257410747622, -0.8794737512064890713909, -0.9863613034027223736025,
      -0.9863613034027223736025, -0.8794737512064890713909, -0.6772815716257410747622,
      -0.4016954246529694575168, -0.08257934547233232460034, 0.2454854871407991489223,
      0.5469481581224268747118, 0.7891405093963935992190, 0.9458172417006346790197};
  static constexpr double s[19] = {
      0.0, 0.3246994692046834874076, 0.6142127126896678174443, 0.8371664782625285748061,
      0.9694002659393304167361, 0.9965844930066698498194, 0.9157733266550574399193,
      0.7357239106731316247742, 0.4759473930370735444314, 0.1645945902807338941437,
      -0.1645945902807338941437, -0.4759473930370735444314, -0.7357239106731316247742,
      -0.9157733266550574399193, -0.9965844930066698498194, -0.9694002659393304167361,
      -0.8371664782625285748061, -0.6142127126896678174443, -0.3246994692046834874076};
};

template <int R, int SIGN, class C>
__host__ __device__ __forceinline__ void yf_dft_odd(C* v) {
  typedef typename wf_real<C>::T Rt;
  constexpr int H = (R - 1) / 2;
  C a[H + 1], b[H + 1];
  C y0 = v[0];
#pragma unroll
  for (int j = 1; j <= H; ++j) {
    a[j] = wf_add(v[j], v[R - j]);
    b[j] = wf_sub(v[j], v[R - j]);
    y0 = wf_add(y0, a[j]);
  }
  const C x0 = v[0];
  v[0] = y0;
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    C re = x0, im = wf_mk((Rt)0, (Rt)0);
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const int t = (j * k) % R;
      const Rt cc = (Rt)OddRoots<R>::c[t], ss = (Rt)OddRoots<R>::s[t];
      re.x += cc * a[j].x;
      re.y += cc * a[j].y;
      im.x += ss * b[j].x;
      im.y += ss * b[j].y;
    }
    const C si = wf_si<SIGN>(im);   // SIGN * i * im
    v[k] = wf_add(re, si);
    v[R - k] = wf_sub(re, si);
  }
}

template <int R, int SIGN, class C>
__host__ __device__ __forceinline__ void yf_dft(C* v) {
  if constexpr (R % 2 == 1 && R >= 5)
    yf_dft_odd<R, SIGN>(v);
  else
    wf_dft<R, SIGN>(v);
}

// one Stockham pass on column w of the tile: inputs x[b + r nb], outputs
// (b/Ns) Ns R + b%Ns + r Ns, twiddle exp(SIGN 2 pi i r (b%Ns) / (Ns R)) read
// from tw[OFF + (r - 1) Ns + b%Ns]
// PT >= 0: each output p >= PNs is multiplied by tw[PT + p - PNs], the
// twiddle of the next pass (a spread pass, whose nb is PNs) at its input p.
template <int R, int SIGN, int N, int Ns, int OFF, int PT = -1, int PNs = 0, class Col, class C>
__device__ __forceinline__ void yf_pass(C* T, Col col, const C* tw) {
  constexpr int nb = N / R;
  constexpr int MB = (nb + 63) / 64;       // butterflies per lane
  const int lane = threadIdx.x & 63;
  C v[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int b = lane + 64 * q;
    if (b < nb) {
      const int k = b % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        C x = T[col(b + r * nb)];
        if (Ns > 1 && r > 0) x = wf_tw<SIGN>(x, tw[OFF + (r - 1) * Ns + k]);
        v[q][r] = x;
      }
    }
  }
  wf_sync();
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int b = lane + 64 * q;
    if (b < nb) {
      const int k = b % Ns;
      yf_dft<R, SIGN>(v[q]);
      const int o = (b / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int p = o + r * Ns;
        C y = v[q][r];
        if (PT >= 0 && p >= PNs) y = wf_tw<SIGN>(y, tw[PT + p - PNs]);
        T[col(p)] = y;
      }
    }
  }
  wf_sync();
}

// The last pass of a plan whose last radix R is a large odd prime (17 and
// up), each butterfly on two lanes.  With a_j = x_j + x_(R-j), b_j = x_j -
// x_(R-j), j = 1..H, H = (R - 1) / 2, output pair k is
//   y_k, y_(R-k) = x_0 + sum_j a_j cos(2 pi j k / R) +- SIGN i sum_j b_j sin(2 pi j k / R).
// Lane h of a butterfly forms the pairs k = 1..K (h = 0) or K + 1..H and
// y_0 (h = 1), K = (H + 1) / 2, walking j once and keeping only its K pair
// sums, not the R inputs and 2H sums of a butterfly per lane (which spill
// the y kernels at R = 19).  Both lanes read each input (an LDS broadcast).
// The multipliers exp(-2 pi i j k / R) sit at tw[N + (h K + k') H + j - 1],
// k' < K the lane's pair, j = 1..H (yf_spread_table).
// In the last Stockham pass (Ns = nb) butterfly b reads and writes the
// same slots b + r nb.  The pass's twiddles are applied by the previous
// pass as it writes (yf_pass with PT >= 0): input b + r nb takes
// tw[OFF + (r - 1) nb + b] = tw[OFF + (b + r nb) - nb].
template <int R, int SIGN, int N, class Col, class C>
__device__ __forceinline__ void yf_pass_spread(C* T, Col col, const C* tw) {
  typedef typename wf_real<C>::T Rt;
  constexpr int H = (R - 1) / 2, K = (H + 1) / 2;
  constexpr int nb = N / R;
  constexpr int NR = (nb + 31) / 32;         // rounds of 32 butterflies
  const int lane = threadIdx.x & 63;
  const int g = lane >> 1, h = lane & 1;
  const C* mt = tw + N + h * K * H;          // this lane's multipliers [k'][j]
  const C zero = wf_mk((Rt)0, (Rt)0);
#pragma unroll
  for (int t = 0; t < NR; ++t) {
    const int b = g + 32 * t;
    const bool act = b < nb;
    C re[K], im[K], sum = zero, x0 = zero;
#pragma unroll
    for (int q = 0; q < K; ++q) re[q] = im[q] = zero;
    if (act) {
      x0 = T[col(b)];
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const C xp = T[col(b + j * nb)], xm = T[col(b + (R - j) * nb)];
        const C aj = wf_add(xp, xm), bj = wf_sub(xp, xm);
        sum = wf_add(sum, aj);
#pragma unroll
        for (int q = 0; q < K; ++q) {
          const C w = mt[q * H + j - 1];        // exp(-2 pi i j k / R)
          re[q].x += w.x * aj.x;
          re[q].y += w.x * aj.y;
          im[q].x -= w.y * bj.x;
          im[q].y -= w.y * bj.y;
        }
      }
    }
    wf_sync();   // the butterflies of this round are read
    if (act) {
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const int k = h * K + q + 1;
        if (k <= H) {
          const C base = wf_add(x0, re[q]), si = wf_si<SIGN>(im[q]);   // SIGN * i * im
          T[col(b + k * nb)] = wf_add(base, si);
          T[col(b + (R - k) * nb)] = wf_sub(base, si);
        }
      }
      if (h == 1) T[col(b)] = wf_add(x0, sum);
    }
  }
  wf_sync();
}
// entries of the multiplier table of a spread pass of radix R
__host__ __device__ constexpr int yf_spread_entries(int R) {
  return 2 * (((R - 1) / 2 + 1) / 2) * ((R - 1) / 2);
}
// the multiplier table of a spread pass of radix R: exp(-2 pi i j k / R) at
// [h][k'][j - 1], k = h K + k' + 1 (zero beyond H)
inline void yf_spread_table(int R, cplx* out) {
  const int H = (R - 1) / 2, K = (H + 1) / 2;
  for (int h = 0; h < 2; ++h)
    for (int q = 0; q < K; ++q)
      for (int j = 1; j <= H; ++j) {
        const int k = h * K + q + 1;
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)((j * k) % R) / R;
        out[(h * K + q) * H + j - 1] = k <= H ? make_double2((double)cosl(a), (double)sinl(a))
                                              : make_double2(0.0, 0.0);
      }
}
// radix of a last pass that runs spread over the lanes
__host__ __device__ constexpr bool yf_spread_radix(int R) { return R >= 17 && R % 2 == 1; }
template <int... Rs> struct YLast;
template <> struct YLast<> {
  static constexpr int value = 0;
};
template <int R> struct YLast<R> {
  static constexpr int value = R;
};
template <int R, int R2, int... Rs> struct YLast<R, R2, Rs...> {
  static constexpr int value = YLast<R2, Rs...>::value;
};

// transform of column w in place (natural order in and out), unnormalised
template <int SIGN, int N, int Ns, int OFF, class Col, class C>
__device__ __forceinline__ void yf_passes(C*, Col, const C*) {}
template <int SIGN, int N, int Ns, int OFF, class Col, class C, int R, int... Rs>
__device__ __forceinline__ void yf_passes(C* T, Col col, const C* tw) {
  constexpr int NOFF = Ns > 1 ? OFF + (R - 1) * Ns : 0;   // the next pass's twiddles
  if constexpr (yf_spread_radix(R)) {
    static_assert(sizeof...(Rs) == 0, "a spread pass runs last");
    yf_pass_spread<R, SIGN, N>(T, col, tw);
  } else if constexpr (sizeof...(Rs) == 1 && yf_spread_radix(YLast<Rs...>::value)) {
    yf_pass<R, SIGN, N, Ns, OFF, NOFF, Ns * R>(T, col, tw);   // twiddles the spread pass's inputs
  } else {
    yf_pass<R, SIGN, N, Ns, OFF>(T, col, tw);
  }
  yf_passes<SIGN, N, Ns * R, NOFF, Col, C, Rs...>(T, col, tw);
}
// element type C: cplx, or cplxf (with an FP32 table tw)
template <int SIGN, class Col, class C, int... Rs>
__device__ __forceinline__ void yf_fft(C* T, Col col, const C* tw) {
  wf_sync();   // the column was just written lane by lane
  yf_passes<SIGN, YPlanT<Rs...>::N, 1, 0, Col, C, Rs...>(T, col, tw);
}

// Direct DFT of column col in place (natural order in and out,
// unnormalised) for any N <= 64 QM: lane l computes X[l + 64 q] as a sum over
// the column (each x_j one LDS broadcast) against rt[t] = exp(-2 pi i t/N).
// Line lengths without a radix plan use it; N^2 multiply-adds per transform
// cost less than the banded sweeps up to a few hundred points.
template <int SIGN, int QM, class Col>
__device__ __forceinline__ void yf_dft_direct(cplx* T, Col col, const cplx* rt, int N) {
  const int lane = threadIdx.x & 63;
  wf_sync();   // the column was just written lane by lane
  cplx acc[QM];
  int idx[QM], stp[QM];
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    acc[q] = mkc(0.0, 0.0);
    idx[q] = 0;
    stp[q] = (lane + 64 * q) % N;
  }
  for (int j = 0; j < N; ++j) {
    const cplx x = T[col(j)];
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const cplx w = rt[idx[q]];
      acc[q] = cadd(acc[q], SIGN < 0 ? cmul(x, w) : cmulc(x, w));
      idx[q] += stp[q];
      if (idx[q] >= N) idx[q] -= N;
    }
  }
  wf_sync();
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int m = lane + 64 * q;
    if (m < N) T[col(m)] = acc[q];
  }
  wf_sync();
}

// Direct DFT of the block's YF_LINES columns on the matrix cores: a
// block-collective form of yf_dft_direct, called by every wave of the block
// with its own column.  With W[m][j] = exp(SIGN 2 pi i m j / N) and the four
// lines' inputs as the columns of B = [Re x_0..3 | Im x_0..3 | 0] (N x 16),
// each wave forms P = Re W . B and Q = Im W . B for its tiles of 16 modes
// with v_mfma_f64_16x16x4f64 (K = 4 inputs a step), and
//   Re X_L = P[., L] - Q[., 4 + L],   Im X_L = P[., 4 + L] + Q[., L].
// Operand layout of the instruction (A 16 x 4, B 4 x 16): lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; result value r of lane l is
// D[4 r + l / 16][l % 16].  W comes from the root table rt[t] = exp(-2 pi i
// t / N) with t = m j mod N advanced by 4 m per step.  Half of each
// instruction's 16 columns carries data (4 lines, re and im).
typedef double yf_v4d __attribute__((ext_vector_type(4)));
template <int SIGN, int QM, class Col>
__device__ __forceinline__ void yf_dft_mfma(cplx* T, Col col, const cplx* rt, int N, bool active) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kq = lane >> 4;
  // which lines take part (a wave may hold its column out)
  __shared__ int act[YF_LINES];
  if (lane == 0) act[w] = active ? 1 : 0;
  __syncthreads();   // every wave's column is written
  const bool take = act[i & 3] != 0;     // line of this lane's B column and output
  const int mt = (N + 15) / 16;                // tiles of 16 modes; this wave takes w, w + 4, ...
  const int ks = (N + 3) / 4;                  // steps of 4 inputs
  yf_v4d P[QM], Q[QM];
  int idx[QM], stp[QM];
#pragma unroll
  for (int t = 0; t < QM; ++t) {
    P[t] = yf_v4d{0.0, 0.0, 0.0, 0.0};
    Q[t] = yf_v4d{0.0, 0.0, 0.0, 0.0};
    const int m = (w + 4 * t) * 16 + i;
    idx[t] = (m * kq) % N;                     // t = m j mod N at j = kq
    stp[t] = (4 * m) % N;
  }
  const Col src = col.peer(i & 3, w, N);       // column of line (n mod 4) for B
  for (int k = 0; k < ks; ++k) {
    const int j = 4 * k + kq;
    double b = 0.0;
    if (i < 8 && j < N && take) {
      const cplx x = T[src(j)];
      b = i < 4 ? x.x : x.y;
    }
#pragma unroll
    for (int t = 0; t < QM; ++t) {
      const int tile = w + 4 * t;
      if (tile < mt) {
        const int m = tile * 16 + i;
        double ar = 0.0, ai = 0.0;
        if (m < N && j < N) {
          const cplx r = rt[idx[t]];
          ar = r.x;
          ai = SIGN < 0 ? r.y : -r.y;
        }
        P[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, b, P[t], 0, 0, 0);
        Q[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, b, Q[t], 0, 0, 0);
        idx[t] += stp[t];
        if (idx[t] >= N) idx[t] -= N;
      }
    }
  }
  __syncthreads();   // every input read before any output lands
#pragma unroll
  for (int t = 0; t < QM; ++t) {
    const int tile = w + 4 * t;
    if (tile < mt) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double pr = P[t][r], qr = Q[t][r];
        const double p4 = __shfl_down(pr, 4, 16), q4 = __shfl_down(qr, 4, 16);
        const int m = tile * 16 + 4 * r + kq;
        if (i < 4 && m < N && take) T[col.peer(i, w, N)(m)] = mkc(pr - q4, p4 + qr);
      }
    }
  }
  __syncthreads();   // the columns of the other waves are written
}

// A column viewed from offset o: element j is element j + o of col.
template <class Col> struct ShiftCol {
  Col c;
  int o;
  __device__ __forceinline__ int operator()(int j) const { return c(j + o); }
};

// DFT of prime length P in place on column col (natural order in and out,
// unnormalised) by Rader's algorithm: with g a primitive root mod P and
// M = P - 1, the outputs X[g^-p] - x[0] are the cyclic convolution over M
// points of u[q] = x[g^q] with v[s] = exp(SIGN 2 pi i g^-s / P), formed as
// IDFT_M(DFT_M(u) . DFT_M(v) / M) with the mixed-radix plan Rs of M; x[0]
// enters as the zero mode of the product, and X[0] = x[0] + DFT_M(u)[0].
// The table tw (4 M entries, staged in LDS): [0, M) the pass twiddles of
// the M-point plan, [M, 2M) DFT_M(v) / M for SIGN = -1, [2M, 3M) the same for
// SIGN = +1, [3M, 4M) (g^q mod P, g^-q mod P) as doubles.  u and the
// convolution live in elements 1..M of the column.
template <int SIGN, int P, class Col, int... Rs>
__device__ __forceinline__ void yf_rader(cplx* T, Col col, const cplx* tw) {
  constexpr int M = P - 1, QM = (M + 63) / 64;
  static_assert(YPlanT<Rs...>::N == M, "the convolution plan has P - 1 points");
  const int lane = threadIdx.x & 63;
  const ShiftCol<Col> sc{col, 1};
  const cplx* perm = tw + 3 * M;
  wf_sync();   // the column was just written lane by lane
  const cplx x0 = T[col(0)];
  cplx u[QM];
#pragma unroll
  for (int t = 0; t < QM; ++t) {
    const int q = lane + 64 * t;
    if (q < M) u[t] = T[col((int)perm[q].x)];
  }
  wf_sync();
#pragma unroll
  for (int t = 0; t < QM; ++t) {
    const int q = lane + 64 * t;
    if (q < M) T[sc(q)] = u[t];
  }
  yf_fft<-1, ShiftCol<Col>, cplx, Rs...>(T, sc, tw);
  const cplx* V = tw + (SIGN < 0 ? M : 2 * M);
  cplx X0 = x0;
#pragma unroll
  for (int t = 0; t < QM; ++t) {
    const int m = lane + 64 * t;
    if (m < M) {
      const cplx U = T[sc(m)];
      cplx W = cmul(U, V[m]);
      if (m == 0) {
        X0 = cadd(x0, U);
        W = cadd(W, x0);
      }
      T[sc(m)] = W;
    }
  }
  yf_fft<1, ShiftCol<Col>, cplx, Rs...>(T, sc, tw);
#pragma unroll
  for (int t = 0; t < QM; ++t) {
    const int p = lane + 64 * t;
    if (p < M) u[t] = T[sc(p)];
  }
  wf_sync();
#pragma unroll
  for (int t = 0; t < QM; ++t) {
    const int p = lane + 64 * t;
    if (p < M) T[col((int)perm[p].y)] = u[t];
  }
  if (lane == 0) T[col(0)] = X0;
  wf_sync();
}

// Plans of the y-line transforms, the template argument of the y kernels.
//   YFFT<Rs...>  mixed-radix Stockham passes, N = prod Rs at compile time,
//                twiddles from the pass-major table;
//   YRader<P, G, Rs...>  prime N = P by Rader's algorithm over the plan Rs
//                of P - 1 (G a primitive root mod P), its table of 4 (P - 1)
//                entries in place of the twiddles;
//   YDirect<QM>  the direct DFT, N = my <= 64 QM at run time, twiddles from
//                the root table exp(-2 pi i t/N).
// tw_len(N): entries of the plan's table that a kernel stages in LDS.
template <int... Rs> struct YFFT {
  static constexpr int Q = (YPlanT<Rs...>::N + 63) / 64;
  static constexpr bool direct = false;
  static constexpr bool f32 = true;   // transforms also on FP32 columns (cplxf)
  static constexpr int NC = YPlanT<Rs...>::N;   // line length (compile time)
  // multiplier rows read from the global table: beyond 448 points the rows
  // staged in LDS would leave one block per CU
  static constexpr bool sym_gl = NC > 448;
  // the pass twiddles (at most N entries), then the multipliers of a spread
  // last pass
  static constexpr int RS =
      yf_spread_radix(YLast<Rs...>::value) ? yf_spread_entries(YLast<Rs...>::value) : 0;
  __host__ __device__ static constexpr int tw_len(int n) { return n + RS; }
  // active = false: this wave's column stays as it is (the call keeps the
  // waves of a block in step for the block-collective plan)
  template <int SIGN, class Col, class C>
  __device__ __forceinline__ static void fft(C* T, Col c, const C* tw, int,
                                             bool active = true) {
    if (active) yf_fft<SIGN, Col, C, Rs...>(T, c, tw);
  }
  template <int SIGN, class Col, class C>   // one wave on its own (the 00 lines)
  __device__ __forceinline__ static void fft_wave(C* T, Col c, const C* tw, int n) {
    fft<SIGN>(T, c, tw, n);
  }
};
template <int P, int G, int... Rs> struct YRader {
  static constexpr int Q = (P + 63) / 64;
  static constexpr bool direct = false;
  static constexpr bool f32 = false;
  static constexpr int NC = P;
  static constexpr bool sym_gl = false;
  __host__ __device__ static constexpr int tw_len(int) { return 4 * (P - 1); }
  template <int SIGN, class Col>
  __device__ __forceinline__ static void fft(cplx* T, Col c, const cplx* tw, int,
                                             bool active = true) {
    if (active) yf_rader<SIGN, P, Col, Rs...>(T, c, tw);
  }
  template <int SIGN, class Col>
  __device__ __forceinline__ static void fft_wave(cplx* T, Col c, const cplx* tw, int n) {
    fft<SIGN>(T, c, tw, n);
  }
};
template <int QM> struct YDirect {
  static constexpr int Q = QM;
  static constexpr bool direct = true;
  static constexpr bool f32 = false;
  static constexpr int NC = 0;                  // line length: my, at run time
  static constexpr bool sym_gl = false;
  __host__ __device__ static constexpr int tw_len(int n) { return n; }
  template <int SIGN, class Col>
  __device__ __forceinline__ static void fft(cplx* T, Col c, const cplx* rt, int N,
                                             bool active = true) {
    if (active) yf_dft_direct<SIGN, QM>(T, c, rt, N);
  }
  template <int SIGN, class Col>
  __device__ __forceinline__ static void fft_wave(cplx* T, Col c, const cplx* rt, int N) {
    yf_dft_direct<SIGN, QM>(T, c, rt, N);
  }
};
// the direct DFT on the matrix cores (block-collective: every wave of a
// line block calls fft at the same point); the 00 lines keep the per-wave sum
template <int QM> struct YMfma {
  static constexpr int Q = QM;
  static constexpr bool direct = true;
  static constexpr bool f32 = false;
  static constexpr int NC = 0;
  static constexpr bool sym_gl = false;
  __host__ __device__ static constexpr int tw_len(int n) { return n; }
  template <int SIGN, class Col>
  __device__ __forceinline__ static void fft(cplx* T, Col c, const cplx* rt, int N,
                                             bool active = true) {
    yf_dft_mfma<SIGN, QM>(T, c, rt, N, active);
  }
  template <int SIGN, class Col>
  __device__ __forceinline__ static void fft_wave(cplx* T, Col c, const cplx* rt, int N) {
    yf_dft_direct<SIGN, QM>(T, c, rt, N);
  }
};

// seven stencil coefficients held in registers
struct C7 {
  double c[7];
};
__device__ __forceinline__ C7 c7_of(const double (&a)[7]) {
  C7 r;
#pragma unroll
  for (int m = 0; m < 7; ++m) r.c[m] = a[m];
  return r;
}

// C(w) = sum_{p=-3..3} c[3+p] exp(i p w) from e = exp(i w)
__device__ __forceinline__ cplx yf_symbol(const C7& k, cplx e1) {
  const double* c = k.c;
  const cplx e2 = cmul(e1, e1), e3 = cmul(e2, e1);
  const double c1 = c[4] + c[2], s1 = c[4] - c[2];
  const double c2 = c[5] + c[1], s2 = c[5] - c[1];
  const double c3 = c[6] + c[0], s3 = c[6] - c[0];
  return mkc(c[3] + c1 * e1.x + c2 * e2.x + c3 * e3.x, s1 * e1.y + s2 * e2.y + s3 * e3.y);
}

// Compact filter of filter_shear_hepta.f90:53-69 (prefilter): symmetric
// hepta LHS p(-3..3) and symmetric 9-point RHS r(-4..4), not scaled by the
// grid step.  Its multiplier is R_f(w) / B_f(w), both real.
__device__ __forceinline__ double yf_filter_symbol(cplx e1) {
  const double p0 = 1.0, p1 = 0.738981897704866, p2 = 0.315003368623046, p3 = 0.053003478448814;
  const double r0 = 0.987412035368317, r1 = 0.749052269410218, r2 = 0.309968182770373,
               r3 = 0.054442102978149, r4 = -0.000179828066167;
  const cplx e2 = cmul(e1, e1), e3 = cmul(e2, e1), e4 = cmul(e2, e2);
  const double B = p0 + 2.0 * (p1 * e1.x + p2 * e2.x + p3 * e3.x);
  const double R = r0 + 2.0 * (r1 * e1.x + r2 * e2.x + r3 * e3.x + r4 * e4.x);
  return R / B;
}

// a / b
__device__ __forceinline__ cplx yf_div(cplx a, cplx b) {
  const double d = 1.0 / (b.x * b.x + b.y * b.y);
  return mkc((a.x * b.x + a.y * b.y) * d, (a.y * b.x - a.x * b.y) * d);
}
