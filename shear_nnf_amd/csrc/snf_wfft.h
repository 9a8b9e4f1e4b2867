// Wave-level FFT of N = 64*R0 points (R0 = 2, 3, 4, 6, 8: N = 128, 192, 256,
// 384, 512).  One wave transforms one sequence held in its registers; the two
// exchanges between stages go through a wave-private LDS region, so no
// block barrier is needed.  Index split N = R0 x 8 x 8:
//
//   lane layout  L: lane l (0..63) holds x[l + 64 m], m = 0..R0-1
//   task layout  T: task t (0..8 R0 - 1; lanes >= 8 R0 idle) holds
//                   x[t + 8 R0 q], q = 0..7
//
// Both layouts read and write natural order with consecutive lanes on
// consecutive elements (coalesced).  l2t: L -> T, t2l: T -> L.
// SIGN = -1: y[k] = sum x[n] exp(-2 pi i nk/N); SIGN = +1: exp(+...).
// Unnormalised.  Twiddles come from rt[j] = exp(-2 pi i j/N), j < N
// (host-computed in extended precision, staged in LDS by the caller).
//
// L -> T, with n = l + 64 m and k = k1 + R0 (q2 + 8 q1), l = l1 + 8 l2:
//   a) per lane: DFT_R0 over m, times w_N^(l k1)            -> E[k1][l]
//   b) task (k1, l1): DFT_8 over l2, times w_64^(l1 q2)     -> F[q2 R0 + k1][l1]
//   c) task t = k1 + R0 q2: DFT_8 over l1                   -> y[t + 8 R0 q1]
// T -> L is the mirror image.
//
// Every stage is a __host__ __device__ function of one lane so that the
// host can run the same index arithmetic lane by lane (unit test).  The
// element type C is cplx (FP64) or cplxf (FP32).
#pragma once
#include "snf_internal.h"

#define WF_EP 72   // row pitch of E (64 + 8: staggers LDS banks)
#define WF_FP 9    // row pitch of F (8 + 1)

// Element types: cplx (FP64) or cplxf (FP32, the x-z stage of the mixed
// precision mode).  wf_real<C>::T is the component type.
template <class C> struct wf_real;
template <> struct wf_real<double2> { typedef double T; };
template <> struct wf_real<float2> { typedef float T; };
__host__ __device__ __forceinline__ double2 wf_mk(double x, double y) { return make_double2(x, y); }
__host__ __device__ __forceinline__ float2 wf_mk(float x, float y) { return make_float2(x, y); }

template <class C>
__host__ __device__ __forceinline__ C wf_add(C a, C b) { return wf_mk(a.x + b.x, a.y + b.y); }
template <class C>
__host__ __device__ __forceinline__ C wf_sub(C a, C b) { return wf_mk(a.x - b.x, a.y - b.y); }
template <class C>
__host__ __device__ __forceinline__ C wf_mul(C a, C b) {
  return wf_mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * w^SIGN-convention: w holds exp(-i theta); SIGN=+1 uses its conjugate
template <int SIGN, class C>
__host__ __device__ __forceinline__ C wf_tw(C a, C w) {
  if (SIGN < 0) return wf_mk(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
  return wf_mk(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}
// SIGN * i * a
template <int SIGN, class C>
__host__ __device__ __forceinline__ C wf_si(C a) {
  return SIGN < 0 ? wf_mk(a.y, -a.x) : wf_mk(-a.y, a.x);
}

template <int R, int SIGN, class C>
__host__ __device__ __forceinline__ void wf_dft(C* v) {
  typedef typename wf_real<C>::T Rt;
  if (R == 2) {
    const C a = v[0], b = v[1];
    v[0] = wf_add(a, b);
    v[1] = wf_sub(a, b);
  } else if (R == 3) {
    const Rt h = (Rt)0.86602540378443864676;   // sqrt(3)/2
    const C t1 = wf_add(v[1], v[2]), t2 = wf_sub(v[1], v[2]);
    const C m = wf_mk(v[0].x - (Rt)0.5 * t1.x, v[0].y - (Rt)0.5 * t1.y);
    const C s = wf_si<SIGN>(wf_mk(h * t2.x, h * t2.y));
    v[0] = wf_add(v[0], t1);
    v[1] = wf_add(m, s);
    v[2] = wf_sub(m, s);
  } else if (R == 4) {
    const C t0 = wf_add(v[0], v[2]), t1 = wf_sub(v[0], v[2]);
    const C t2 = wf_add(v[1], v[3]), t3 = wf_si<SIGN>(wf_sub(v[1], v[3]));
    v[0] = wf_add(t0, t2);
    v[2] = wf_sub(t0, t2);
    v[1] = wf_add(t1, t3);
    v[3] = wf_sub(t1, t3);
  } else if (R == 6) {
    // 6 = 2 x 3: E_b = DFT_3(x[b], x[b+2], x[b+4]); y[k] = E0[k] + w6^k E1[k],
    // y[k+3] = E0[k] - w6^k E1[k]
    C e[3] = {v[0], v[2], v[4]}, o[3] = {v[1], v[3], v[5]};
    wf_dft<3, SIGN>(e);
    wf_dft<3, SIGN>(o);
    const Rt h = (Rt)0.86602540378443864676;
    const C w1 = wf_mk((Rt)0.5, (Rt)SIGN * h), w2 = wf_mk((Rt)-0.5, (Rt)SIGN * h);
    const C o1 = wf_mul(o[1], w1), o2 = wf_mul(o[2], w2);
    v[0] = wf_add(e[0], o[0]);
    v[3] = wf_sub(e[0], o[0]);
    v[1] = wf_add(e[1], o1);
    v[4] = wf_sub(e[1], o1);
    v[2] = wf_add(e[2], o2);
    v[5] = wf_sub(e[2], o2);
  } else if (R == 8) {
    C e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    wf_dft<4, SIGN>(e);
    wf_dft<4, SIGN>(o);
    const Rt r = (Rt)0.70710678118654752440;   // sqrt(2)/2
    // w8 = r (1 + SIGN i), w8^2 = SIGN i, w8^3 = r (-1 + SIGN i)
    const C o1 = wf_mk(r * (o[1].x - (Rt)SIGN * o[1].y), r * (o[1].y + (Rt)SIGN * o[1].x));
    const C o2 = wf_si<SIGN>(o[2]);
    const C o3 = wf_mk(r * (-o[3].x - (Rt)SIGN * o[3].y), r * (-o[3].y + (Rt)SIGN * o[3].x));
    v[0] = wf_add(e[0], o[0]);
    v[4] = wf_sub(e[0], o[0]);
    v[1] = wf_add(e[1], o1);
    v[5] = wf_sub(e[1], o1);
    v[2] = wf_add(e[2], o2);
    v[6] = wf_sub(e[2], o2);
    v[3] = wf_add(e[3], o3);
    v[7] = wf_sub(e[3], o3);
  }
}

template <int R0, int SIGN, class C = cplx>
struct WFFT {
  static constexpr int N = 64 * R0;
  static constexpr int NT = 8 * R0;          // active tasks
  static constexpr int LDS = 8 * R0 * WF_FP > R0 * WF_EP ? 8 * R0 * WF_FP : R0 * WF_EP;

  // ---- L -> T ----
  __host__ __device__ static void l2t_a(int l, C* a, C* E, const C* rt) {
    wf_dft<R0, SIGN>(a);
    E[l] = a[0];
#pragma unroll
    for (int k1 = 1; k1 < R0; ++k1) E[k1 * WF_EP + l] = wf_tw<SIGN>(a[k1], rt[l * k1]);
  }
  // task t' = l1 + 8 k1; reads E, the caller separates reads from writes
  __host__ __device__ static void l2t_b_read(int t, const C* E, C* d) {
    const int l1 = t & 7, k1 = t >> 3;
#pragma unroll
    for (int l2 = 0; l2 < 8; ++l2) d[l2] = E[k1 * WF_EP + l1 + 8 * l2];
  }
  __host__ __device__ static void l2t_b_write(int t, C* d, C* F, const C* rt) {
    const int l1 = t & 7, k1 = t >> 3;
    wf_dft<8, SIGN>(d);
    F[(0 * R0 + k1) * WF_FP + l1] = d[0];
#pragma unroll
    for (int q2 = 1; q2 < 8; ++q2)
      F[(q2 * R0 + k1) * WF_FP + l1] = wf_tw<SIGN>(d[q2], rt[l1 * q2 * R0]);
  }
  // task t = k1 + R0 q2: row t of F -> y[t + 8 R0 q1]
  __host__ __device__ static void l2t_c(int t, const C* F, C* out) {
#pragma unroll
    for (int l1 = 0; l1 < 8; ++l1) out[l1] = F[t * WF_FP + l1];
    wf_dft<8, SIGN>(out);
  }

  // ---- T -> L ----
  // task t = n1 + R0 q2 holds v[q1] = x[t + 8 R0 q1]; rows of F are
  // (l1, n1) with n1 fastest, which keeps these writes on distinct banks
  __host__ __device__ static void t2l_a(int t, C* v, C* F, const C* rt) {
    const int n1 = t % R0, q2 = t / R0;
    wf_dft<8, SIGN>(v);
    F[(0 * R0 + n1) * WF_FP + q2] = v[0];
#pragma unroll
    for (int l1 = 1; l1 < 8; ++l1)
      F[(l1 * R0 + n1) * WF_FP + q2] = wf_tw<SIGN>(v[l1], rt[q2 * l1 * R0]);
  }
  // task t' = l1 + 8 n1 reads row (l1, n1)
  __host__ __device__ static void t2l_b_read(int t, const C* F, C* u) {
    const int row = (t & 7) * R0 + (t >> 3);
#pragma unroll
    for (int q2 = 0; q2 < 8; ++q2) u[q2] = F[row * WF_FP + q2];
  }
  __host__ __device__ static void t2l_b_write(int t, C* u, C* E) {
    const int l1 = t & 7, n1 = t >> 3;
    wf_dft<8, SIGN>(u);
#pragma unroll
    for (int l2 = 0; l2 < 8; ++l2) E[n1 * WF_EP + l1 + 8 * l2] = u[l2];
  }
  // lane l -> a[m] = X[l + 64 m]
  __host__ __device__ static void t2l_c(int l, const C* E, C* a, const C* rt) {
    a[0] = E[l];
#pragma unroll
    for (int n1 = 1; n1 < R0; ++n1) a[n1] = wf_tw<SIGN>(E[n1 * WF_EP + l], rt[n1 * l]);
    wf_dft<R0, SIGN>(a);
  }
};

// order the wave's LDS accesses across lanes (LDS executes a wave's
// instructions in order; this keeps the compiler from reordering them)
__device__ __forceinline__ void wf_sync() {
#ifdef __HIP_DEVICE_COMPILE__
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#endif
}

// Device drivers: a[R0] (lane layout) -> out[8] (task layout) and back.
// buf: this wave's LDS region of WFFT<R0,SIGN>::LDS values; rt: N roots in LDS.
template <int R0, int SIGN, class C>
__device__ __forceinline__ void wfft_l2t(C* a, C* out, C* buf, const C* rt) {
  typedef WFFT<R0, SIGN, C> W;
  const int lane = threadIdx.x & 63;
  wf_sync();
  W::l2t_a(lane, a, buf, rt);
  wf_sync();
  C d[8];
  if (lane < W::NT) W::l2t_b_read(lane, buf, d);
  wf_sync();
  if (lane < W::NT) W::l2t_b_write(lane, d, buf, rt);
  wf_sync();
  if (lane < W::NT) W::l2t_c(lane, buf, out);
}

template <int R0, int SIGN, class C>
__device__ __forceinline__ void wfft_t2l(C* v, C* a, C* buf, const C* rt) {
  typedef WFFT<R0, SIGN, C> W;
  const int lane = threadIdx.x & 63;
  wf_sync();
  if (lane < W::NT) W::t2l_a(lane, v, buf, rt);
  wf_sync();
  C u[8];
  if (lane < W::NT) W::t2l_b_read(lane, buf, u);
  wf_sync();
  if (lane < W::NT) W::t2l_b_write(lane, u, buf);
  wf_sync();
  W::t2l_c(lane, buf, a, rt);
}

// Two independent transforms of one wave, stage by stage: each exchange
// waits for the LDS once for both sequences, which halves the exposed LDS
// latency of a kernel that runs at few waves per SIMD.  buf0, buf1: two
// disjoint regions of WFFT<R0,SIGN>::LDS values.
template <int R0, int SIGN, class C>
__device__ __forceinline__ void wfft_l2t2(C* a0, C* a1, C* out0, C* out1, C* buf0, C* buf1,
                                          const C* rt) {
  typedef WFFT<R0, SIGN, C> W;
  const int lane = threadIdx.x & 63;
  wf_sync();
  W::l2t_a(lane, a0, buf0, rt);
  W::l2t_a(lane, a1, buf1, rt);
  wf_sync();
  C d0[8], d1[8];
  if (lane < W::NT) {
    W::l2t_b_read(lane, buf0, d0);
    W::l2t_b_read(lane, buf1, d1);
  }
  wf_sync();
  if (lane < W::NT) {
    W::l2t_b_write(lane, d0, buf0, rt);
    W::l2t_b_write(lane, d1, buf1, rt);
  }
  wf_sync();
  if (lane < W::NT) {
    W::l2t_c(lane, buf0, out0);
    W::l2t_c(lane, buf1, out1);
  }
}

template <int R0, int SIGN, class C>
__device__ __forceinline__ void wfft_t2l2(C* v0, C* v1, C* a0, C* a1, C* buf0, C* buf1,
                                          const C* rt) {
  typedef WFFT<R0, SIGN, C> W;
  const int lane = threadIdx.x & 63;
  wf_sync();
  if (lane < W::NT) {
    W::t2l_a(lane, v0, buf0, rt);
    W::t2l_a(lane, v1, buf1, rt);
  }
  wf_sync();
  C u0[8], u1[8];
  if (lane < W::NT) {
    W::t2l_b_read(lane, buf0, u0);
    W::t2l_b_read(lane, buf1, u1);
  }
  wf_sync();
  if (lane < W::NT) {
    W::t2l_b_write(lane, u0, buf0);
    W::t2l_b_write(lane, u1, buf1);
  }
  wf_sync();
  W::t2l_c(lane, buf0, a0, rt);
  W::t2l_c(lane, buf1, a1, rt);
}
