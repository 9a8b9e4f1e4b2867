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

// ---- compile-time plans for the small sizes of the ensemble boxes ----
// The same radix order as the host's plan_of (4s, then 2s, then 3s); with n
// and the stride Ns constants, the index arithmetic of a pass (t / nb,
// j % Ns, j / Ns) compiles to multiplies and shifts instead of the integer
// divisions that dominate the runtime-plan passes at these sizes.  The
// arithmetic on the data is the same, so the results are the same bits.
__host__ __device__ constexpr int ct_nr(int n) {
  int c = 0, m = n;
  while (m % 4 == 0 && m > 1) { m /= 4; ++c; }
  while (m % 2 == 0 && m > 1) { m /= 2; ++c; }
  while (m % 3 == 0 && m > 1) { m /= 3; ++c; }
  return c;
}
__host__ __device__ constexpr int ct_r(int n, int p) {
  int c = 0, m = n;
  while (m % 4 == 0 && m > 1) { if (c == p) return 4; m /= 4; ++c; }
  while (m % 2 == 0 && m > 1) { if (c == p) return 2; m /= 2; ++c; }
  while (m % 3 == 0 && m > 1) { if (c == p) return 3; m /= 3; ++c; }
  return 1;
}
template <int R, int MB, int N, int NS, class C>
__device__ __forceinline__ void stockham_pass_ip_ct(C* buf, int nseq, const cplx* tw, int sign) {
  constexpr int nb = N / R, stw = N / (NS * R);
  const int total = nseq * nb;
  C v[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int t = threadIdx.x + q * blockDim.x;
    if (t < total) {
      const int s = t / nb;
      const int j = t - s * nb;
      const int k = j % NS;
      const C* sp = buf + s * N;
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
      const int k = j % NS;
      dft_small<R>(v[q], sign);
      C* dp = buf + s * N + (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dp[r * NS] = v[q][r];
    }
  }
  __syncthreads();
}
template <int MAXV, int N, int P, int NS, class C>
__device__ __forceinline__ void fft_ip_ct(C* buf, int nseq, const cplx* tw, int sign) {
  if constexpr (P < ct_nr(N)) {
    constexpr int R = ct_r(N, P);
    stockham_pass_ip_ct<R, MAXV / R, N, NS>(buf, nseq, tw, sign);
    fft_ip_ct<MAXV, N, P + 1, NS * R>(buf, nseq, tw, sign);
  }
}
template <int R, int N, int NS, class C>
__device__ __forceinline__ void stockham_pass_ct(const C* src, C* dst, int nseq, const cplx* tw,
                                                 int sign, int tid, int nthr) {
  constexpr int nb = N / R, stw = N / (NS * R);
  const int total = nseq * nb;
  for (int t = tid; t < total; t += nthr) {
    const int s = t / nb;
    const int j = t - s * nb;
    const int k = j % NS;
    const C* sp = src + s * N;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[r] = sp[j + r * nb];
      if (r > 0) v[r] = cmul_sign(v[r], tw[r * k * stw], sign);
    }
    dft_small<R>(v, sign);
    C* dp = dst + s * N + (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dp[r * NS] = v[r];
  }
}
// ping-pong passes P.. of size N: returns the buffer that holds the result
template <int N, int P, int NS, class C>
__device__ __forceinline__ C* fft_pp_ct(C* src, C* dst, int nseq, const cplx* tw, int sign) {
  if constexpr (P < ct_nr(N)) {
    constexpr int R = ct_r(N, P);
    stockham_pass_ct<R, N, NS>(src, dst, nseq, tw, sign, threadIdx.x, blockDim.x);
    __syncthreads();
    return fft_pp_ct<N, P + 1, NS * R>(dst, src, nseq, tw, sign);
  } else {
    return src;
  }
}
// sizes with a compile-time plan: the ensemble boxes' 32, 48, 64 and 96
#define SNF_FFT_CT_CASES(F) F(32) F(48) F(64) F(96)

// MAXV: complex values a thread may hold per pass (R * butterflies)
template <int MAXV, class C>
__device__ __forceinline__ void fft_lds_ip(C* buf, int nseq, const FFTPlan& pl,
                                           const cplx* tw, int sign) {
  switch (pl.n) {
#define SNF_FFT_CT_IP(n)                                       \
    case n:                                                     \
      fft_ip_ct<MAXV, n, 0, 1>(buf, nseq, tw, sign);            \
      return;
    SNF_FFT_CT_CASES(SNF_FFT_CT_IP)
#undef SNF_FFT_CT_IP
    default:
      break;
  }
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
// CT = false: the runtime plan at every size (the spectral-to-hybrid kernel
// keeps it: its compile-time passes raise its registers from 6 waves per
// SIMD to 4)
template <bool CT = true, class C>
__device__ __forceinline__ C* fft_lds(C* a, C* b, int nseq, const FFTPlan& pl,
                                      const cplx* tw, int sign) {
  if (CT) switch (pl.n) {
#define SNF_FFT_CT_PP(n) \
    case n:               \
      return fft_pp_ct<n, 0, 1>(a, b, nseq, tw, sign);
    SNF_FFT_CT_CASES(SNF_FFT_CT_PP)
#undef SNF_FFT_CT_PP
    default:
      break;
  }
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
