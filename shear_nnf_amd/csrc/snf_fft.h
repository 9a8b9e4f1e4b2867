// Mixed-radix (4, 2, 3) Stockham FFT on sequences resident in LDS.
//
// Used for the x and z transforms of the nonlinear stage (the role of the
// FFTW 2-D plans in fou3D_thread.f:1-73).  All sizes of the supported grids
// are 2^a 3^b.  Sequences are contiguous in LDS; a pass reads from one
// buffer and writes the other (ping-pong), so the caller provides 2 buffers
// of nseq*n complex values and gets back the one that holds the result.
// Twiddles come from a table tw[t] = exp(-2 pi i t / n) computed on the
// host in extended precision and rounded once to binary64.
#pragma once
#include "snf_internal.h"

struct FFTPlan {
  int n, nr;
  int r[12];
};

__device__ __forceinline__ cplx cmul_sign(cplx a, cplx w, int sign) {
  // a * w (sign<0) or a * conj(w) (sign>0)
  return sign < 0 ? cmul(a, w) : cmulc(a, w);
}

template <int R>
__device__ __forceinline__ void dft_small(cplx* v, int sign) {
  if (R == 2) {
    cplx a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 4) {
    cplx t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cplx t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    // t3 = d * (sign * i)
    cplx t3 = sign < 0 ? mkc(d.y, -d.x) : mkc(-d.y, d.x);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else if (R == 3) {
    const double h = 0.86602540378443864676;  // sqrt(3)/2
    cplx t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    cplx m = mkc(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    // s = sign * h * i * t2
    cplx sx = sign < 0 ? mkc(h * t2.y, -h * t2.x) : mkc(-h * t2.y, h * t2.x);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, sx);
    v[2] = csub(m, sx);
  }
}

template <int R>
__device__ __forceinline__ void stockham_pass(const cplx* src, cplx* dst, int n, int nseq,
                                              int Ns, const cplx* tw, int sign, int tid,
                                              int nthr) {
  const int nb = n / R;
  const int total = nseq * nb;
  const int stw = n / (Ns * R);
  for (int t = tid; t < total; t += nthr) {
    const int s = t / nb;
    const int j = t - s * nb;
    const int k = j % Ns;
    const cplx* sp = src + s * n;
    cplx v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[r] = sp[j + r * nb];
      if (r > 0) v[r] = cmul_sign(v[r], tw[r * k * stw], sign);
    }
    dft_small<R>(v, sign);
    cplx* dp = dst + s * n + (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dp[r * Ns] = v[r];
  }
}

// In-place variant: every pass first loads all butterfly inputs of the
// thread into registers (at most MB butterflies per thread), synchronises,
// then writes the outputs to their Stockham positions in the same buffer.
// Halves the LDS footprint of the ping-pong form.  The caller guarantees
// nseq*n/R <= MB*blockDim.x for every radix R of the plan.
template <int R, int MB>
__device__ __forceinline__ void stockham_pass_ip(cplx* buf, int n, int nseq, int Ns,
                                                 const cplx* tw, int sign) {
  const int nb = n / R;
  const int total = nseq * nb;
  const int stw = n / (Ns * R);
  cplx v[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int t = threadIdx.x + q * blockDim.x;
    if (t < total) {
      const int s = t / nb;
      const int j = t - s * nb;
      const int k = j % Ns;
      const cplx* sp = buf + s * n;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cplx x = sp[j + r * nb];
        if (r > 0) x = cmul_sign(x, tw[r * k * stw], sign);
        v[q][r] = x;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int t = threadIdx.x + q * blockDim.x;
    if (t < total) {
      const int s = t / nb;
      const int j = t - s * nb;
      const int k = j % Ns;
      dft_small<R>(v[q], sign);
      cplx* dp = buf + s * n + (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dp[r * Ns] = v[q][r];
    }
  }
  __syncthreads();
}

// MAXV: complex values a thread may hold per pass (R * butterflies)
template <int MAXV>
__device__ __forceinline__ void fft_lds_ip(cplx* buf, int nseq, const FFTPlan& pl,
                                           const cplx* tw, int sign) {
  int Ns = 1;
  for (int p = 0; p < pl.nr; ++p) {
    const int R = pl.r[p];
    if (R == 4)
      stockham_pass_ip<4, MAXV / 4>(buf, pl.n, nseq, Ns, tw, sign);
    else if (R == 2)
      stockham_pass_ip<2, MAXV / 2>(buf, pl.n, nseq, Ns, tw, sign);
    else
      stockham_pass_ip<3, MAXV / 3>(buf, pl.n, nseq, Ns, tw, sign);
    Ns *= R;
  }
}

// largest number of butterflies per thread any pass of the plan needs
inline int fft_max_values(const FFTPlan& pl, int nseq, int nthreads) {
  int m = 0;
  for (int p = 0; p < pl.nr; ++p) {
    const int R = pl.r[p];
    const int total = nseq * (pl.n / R);
    const int per = (total + nthreads - 1) / nthreads;
    if (per * R > m) m = per * R;
  }
  return m;
}

// Transforms nseq sequences in a (result in the returned buffer, a or b).
// sign = -1: forward exp(-i...), sign = +1: inverse exp(+i...), unnormalised.
__device__ __forceinline__ cplx* fft_lds(cplx* a, cplx* b, int nseq, const FFTPlan& pl,
                                         const cplx* tw, int sign) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  cplx* src = a;
  cplx* dst = b;
  int Ns = 1;
  for (int p = 0; p < pl.nr; ++p) {
    const int R = pl.r[p];
    if (R == 4)
      stockham_pass<4>(src, dst, pl.n, nseq, Ns, tw, sign, tid, nthr);
    else if (R == 2)
      stockham_pass<2>(src, dst, pl.n, nseq, Ns, tw, sign, tid, nthr);
    else
      stockham_pass<3>(src, dst, pl.n, nseq, Ns, tw, sign, tid, nthr);
    __syncthreads();
    cplx* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}
