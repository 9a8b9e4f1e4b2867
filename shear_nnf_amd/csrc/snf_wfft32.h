// Half-wave FFT of N = 32*R0 points (R0 = 2, 3: N = 64, 96): each 32-lane
// half of a wave transforms its own sequence in registers, with two
// exchanges through a region of LDS private to the half.  The same scheme as
// WFFT (snf_wfft.h) with the index split N = R0 x 8 x 4:
//
//   lane layout  L: lane l (0..31 of its half) holds x[l + 32 m], m < R0
//   task layout  T: task t (0..4 R0 - 1; lanes >= 4 R0 idle) holds
//                   x[t + 4 R0 q], q = 0..7
//
// L -> T, with n = l + 32 m, l = l1 + 8 l2 (l2 < 4), k = k1 + R0 (q2 + 4 q1):
//   a) per lane: DFT_R0 over m, times w_N^(l k1)            -> E[k1][l]
//   b) task (k1, l1): DFT_4 over l2, times w_32^(l1 q2)     -> F[q2 R0 + k1][l1]
//   c) task t = k1 + R0 q2: DFT_8 over l1                   -> y[t + 4 R0 q1]
// T -> L is the mirror image:
//   a') task t = n1 + R0 q2 (x[t + 4 R0 q1]): DFT_8 over q1 -> index l1, times
//       w_32^(q2 l1)                                        -> F[l1 R0 + n1][q2]
//   b') task (l1, n1): DFT_4 over q2                        -> E[n1][l1 + 8 l2]
//   c') lane l: times w_N^(n1 l), DFT_R0 over n1            -> X[l + 32 m]
// SIGN and the root table rt[j] = exp(-2 pi i j/N) as in WFFT; unnormalised.
// Every stage is a __host__ __device__ function of one lane (l < 32), so the
// host runs the same index arithmetic (tests).
#pragma once
#include "snf_wfft.h"

#define WF32_EP 40   // row pitch of E (32 + 8)
#define WF32_FP 9    // row pitch of F in L -> T (8 + 1)
#define WF32_GP 5    // row pitch of F in T -> L (4 + 1)

template <int R0, int SIGN, class C = cplx>
struct WFFT32 {
  static constexpr int N = 32 * R0;
  static constexpr int NT = 4 * R0;   // active tasks of the task layout
  static constexpr int NB = 8 * R0;   // tasks of the middle stage
  static constexpr int LDS = R0 * WF32_EP > 4 * R0 * WF32_FP
                                 ? (R0 * WF32_EP > 8 * R0 * WF32_GP ? R0 * WF32_EP : 8 * R0 * WF32_GP)
                                 : (4 * R0 * WF32_FP > 8 * R0 * WF32_GP ? 4 * R0 * WF32_FP
                                                                          : 8 * R0 * WF32_GP);

  // ---- L -> T ----
  __host__ __device__ static void l2t_a(int l, C* a, C* E, const C* rt) {
    wf_dft<R0, SIGN>(a);
    E[l] = a[0];
#pragma unroll
    for (int k1 = 1; k1 < R0; ++k1) E[k1 * WF32_EP + l] = wf_tw<SIGN>(a[k1], rt[l * k1]);
  }
  // task t' = l1 + 8 k1 (t' < NB)
  __host__ __device__ static void l2t_b_read(int t, const C* E, C* d) {
    const int l1 = t & 7, k1 = t >> 3;
#pragma unroll
    for (int l2 = 0; l2 < 4; ++l2) d[l2] = E[k1 * WF32_EP + l1 + 8 * l2];
  }
  __host__ __device__ static void l2t_b_write(int t, C* d, C* F, const C* rt) {
    const int l1 = t & 7, k1 = t >> 3;
    wf_dft<4, SIGN>(d);
    F[k1 * WF32_FP + l1] = d[0];
#pragma unroll
    for (int q2 = 1; q2 < 4; ++q2)
      F[(q2 * R0 + k1) * WF32_FP + l1] = wf_tw<SIGN>(d[q2], rt[l1 * q2 * R0]);
  }
  // task t = k1 + R0 q2 (t < NT): row t of F -> y[t + 4 R0 q1]
  __host__ __device__ static void l2t_c(int t, const C* F, C* out) {
#pragma unroll
    for (int l1 = 0; l1 < 8; ++l1) out[l1] = F[t * WF32_FP + l1];
    wf_dft<8, SIGN>(out);
  }

  // ---- T -> L ----
  // task t = n1 + R0 q2 holds v[q1] = x[t + 4 R0 q1]
  __host__ __device__ static void t2l_a(int t, C* v, C* F, const C* rt) {
    const int n1 = t % R0, q2 = t / R0;
    wf_dft<8, SIGN>(v);
    F[n1 * WF32_GP + q2] = v[0];
#pragma unroll
    for (int l1 = 1; l1 < 8; ++l1)
      F[(l1 * R0 + n1) * WF32_GP + q2] = wf_tw<SIGN>(v[l1], rt[q2 * l1 * R0]);
  }
  // task t' = l1 + 8 n1 (t' < NB) reads row (l1, n1)
  __host__ __device__ static void t2l_b_read(int t, const C* F, C* u) {
    const int row = (t & 7) * R0 + (t >> 3);
#pragma unroll
    for (int q2 = 0; q2 < 4; ++q2) u[q2] = F[row * WF32_GP + q2];
  }
  __host__ __device__ static void t2l_b_write(int t, C* u, C* E) {
    const int l1 = t & 7, n1 = t >> 3;
    wf_dft<4, SIGN>(u);
#pragma unroll
    for (int l2 = 0; l2 < 4; ++l2) E[n1 * WF32_EP + l1 + 8 * l2] = u[l2];
  }
  // lane l -> a[m] = X[l + 32 m]
  __host__ __device__ static void t2l_c(int l, const C* E, C* a, const C* rt) {
    a[0] = E[l];
#pragma unroll
    for (int n1 = 1; n1 < R0; ++n1) a[n1] = wf_tw<SIGN>(E[n1 * WF32_EP + l], rt[n1 * l]);
    wf_dft<R0, SIGN>(a);
  }
};

// Device drivers for one half-wave: h = lane / 32 selects the half's LDS
// region buf (WFFT32::LDS values); l = lane % 32.
template <int R0, int SIGN, class C>
__device__ __forceinline__ void wfft32_l2t(C* a, C* out, C* buf, const C* rt) {
  typedef WFFT32<R0, SIGN, C> W;
  const int l = threadIdx.x & 31;
  wf_sync();
  W::l2t_a(l, a, buf, rt);
  wf_sync();
  C d[4];
  if (l < W::NB) W::l2t_b_read(l, buf, d);
  wf_sync();
  if (l < W::NB) W::l2t_b_write(l, d, buf, rt);
  wf_sync();
  if (l < W::NT) W::l2t_c(l, buf, out);
}
template <int R0, int SIGN, class C>
__device__ __forceinline__ void wfft32_t2l(C* v, C* a, C* buf, const C* rt) {
  typedef WFFT32<R0, SIGN, C> W;
  const int l = threadIdx.x & 31;
  wf_sync();
  if (l < W::NT) W::t2l_a(l, v, buf, rt);
  wf_sync();
  C u[4];
  if (l < W::NB) W::t2l_b_read(l, buf, u);
  wf_sync();
  if (l < W::NB) W::t2l_b_write(l, u, buf);
  wf_sync();
  W::t2l_c(l, buf, a, rt);
}
