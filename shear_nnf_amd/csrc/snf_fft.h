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
#include "snf_wfft.h"

struct FFTPlan {
  int n, nr;
  int r[12];
};

// The passes are templates on the element type C: cplx (FP64), or cplxf
// for the x-z stage of the mixed precision mode.  Twiddles stay FP64 in
// global memory and are rounded to C when read.
__device__ __forceinline__ cplx cmul_sign(cplx a, cplx w, int sign) {
  // a * w (sign<0) or a * conj(w) (sign>0)
  return sign < 0 ? cmul(a, w) : cmulc(a, w);
}
__device__ __forceinline__ cplxf cmul_sign(cplxf a, cplx wd, int sign) {
  const cplxf w = to_h<cplxf>(wd);
  return sign < 0 ? make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x)
                  : make_float2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}

template <int R, class C>
__device__ __forceinline__ void dft_small(C* v, int sign) {
  typedef typename wf_real<C>::T T;
  if (R == 2) {
    C a = v[0], b = v[1];
    v[0] = wf_add(a, b);
    v[1] = wf_sub(a, b);
  } else if (R == 4) {
    C t0 = wf_add(v[0], v[2]), t1 = wf_sub(v[0], v[2]);
    C t2 = wf_add(v[1], v[3]), d = wf_sub(v[1], v[3]);
    // t3 = d * (sign * i)
    C t3 = sign < 0 ? wf_mk(d.y, -d.x) : wf_mk(-d.y, d.x);
    v[0] = wf_add(t0, t2);
    v[2] = wf_sub(t0, t2);
    v[1] = wf_add(t1, t3);
    v[3] = wf_sub(t1, t3);
  } else if (R == 3) {
    const T h = (T)0.86602540378443864676;  // sqrt(3)/2
    C t1 = wf_add(v[1], v[2]), t2 = wf_sub(v[1], v[2]);
    C m = wf_mk(v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y);
    // s = sign * h * i * t2
    C sx = sign < 0 ? wf_mk(h * t2.y, -h * t2.x) : wf_mk(-h * t2.y, h * t2.x);
    v[0] = wf_add(v[0], t1);
    v[1] = wf_add(m, sx);
    v[2] = wf_sub(m, sx);
  }
}

template <int R, class C>
__device__ __forceinline__ void stockham_pass(const C* src, C* dst, int n, int nseq,
                                              int Ns, const cplx* tw, int sign, int tid,
                                              int nthr) {
  const int nb = n / R;
  const int total = nseq * nb;
  const int stw = n / (Ns * R);
  for (int t = tid; t < total; t += nthr) {
    const int s = t / nb;
    const int j = t - s * nb;
    const int k = j % Ns;
    const C* sp = src + s * n;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[r] = sp[j + r * nb];
      if (r > 0) v[r] = cmul_sign(v[r], tw[r * k * stw], sign);
    }
    dft_small<R>(v, sign);
    C* dp = dst + s * n + (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dp[r * Ns] = v[r];
  }
}

// In-place variant: every pass first loads all butterfly inputs of the
// thread into registers (at most MB butterflies per thread), synchronises,
// then writes the outputs to their Stockham positions in the same buffer.
// Halves the LDS footprint of the ping-pong form.  The caller guarantees
// nseq*n/R <= MB*blockDim.x for every radix R of the plan.
template <int R, int MB, class C>
__device__ __forceinline__ void stockham_pass_ip(C* buf, int n, int nseq, int Ns,
                                                 const cplx* tw, int sign) {
  const int nb = n / R;
  const int total = nseq * nb;
  const int stw = n / (Ns * R);
  C v[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int t = threadIdx.x + q * blockDim.x;
    if (t < total) {
      const int s = t / nb;
      const int j = t - s * nb;
      const int k = j % Ns;
      const C* sp = buf + s * n;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        C x = sp[j + r * nb];
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
      C* dp = buf + s * n + (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dp[r * Ns] = v[q][r];
    }
  }
  __syncthreads();
}

// MAXV: complex values a thread may hold per pass (R * butterflies)
template <int MAXV, class C>
__device__ __forceinline__ void fft_lds_ip(C* buf, int nseq, const FFTPlan& pl,
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
template <class C>
__device__ __forceinline__ C* fft_lds(C* a, C* b, int nseq, const FFTPlan& pl,
                                      const cplx* tw, int sign) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  C* src = a;
  C* dst = b;
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
    C* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}
