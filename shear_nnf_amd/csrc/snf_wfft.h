// Wave-level FFT of N = 64*R0 points (R0 = 2, 3, 4, 6: N = 128, 192, 256,
// 384).  One wave transforms one sequence held in its registers; the two
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
// host can run the same index arithmetic lane by lane (unit test).
#pragma once
#include "snf_internal.h"

#define WF_EP 72   // row pitch of E (64 + 8: staggers LDS banks)
#define WF_FP 9    // row pitch of F (8 + 1)

__host__ __device__ __forceinline__ cplx wf_add(cplx a, cplx b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ cplx wf_sub(cplx a, cplx b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ cplx wf_mul(cplx a, cplx b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * w^SIGN-convention: w holds exp(-i theta); SIGN=+1 uses its conjugate
template <int SIGN>
__host__ __device__ __forceinline__ cplx wf_tw(cplx a, cplx w) {
  if (SIGN < 0) return make_double2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
  return make_double2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}
// SIGN * i * a
template <int SIGN>
__host__ __device__ __forceinline__ cplx wf_si(cplx a) {
  return SIGN < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

template <int R, int SIGN>
__host__ __device__ __forceinline__ void wf_dft(cplx* v) {
  if (R == 2) {
    const cplx a = v[0], b = v[1];
    v[0] = wf_add(a, b);
    v[1] = wf_sub(a, b);
  } else if (R == 3) {
    const double h = 0.86602540378443864676;   // sqrt(3)/2
    const cplx t1 = wf_add(v[1], v[2]), t2 = wf_sub(v[1], v[2]);
    const cplx m = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    const cplx s = wf_si<SIGN>(make_double2(h * t2.x, h * t2.y));
    v[0] = wf_add(v[0], t1);
    v[1] = wf_add(m, s);
    v[2] = wf_sub(m, s);
  } else if (R == 4) {
    const cplx t0 = wf_add(v[0], v[2]), t1 = wf_sub(v[0], v[2]);
    const cplx t2 = wf_add(v[1], v[3]), t3 = wf_si<SIGN>(wf_sub(v[1], v[3]));
    v[0] = wf_add(t0, t2);
    v[2] = wf_sub(t0, t2);
    v[1] = wf_add(t1, t3);
    v[3] = wf_sub(t1, t3);
  } else if (R == 6) {
    // 6 = 2 x 3: E_b = DFT_3(x[b], x[b+2], x[b+4]); y[k] = E0[k] + w6^k E1[k],
    // y[k+3] = E0[k] - w6^k E1[k]
    cplx e[3] = {v[0], v[2], v[4]}, o[3] = {v[1], v[3], v[5]};
    wf_dft<3, SIGN>(e);
    wf_dft<3, SIGN>(o);
    const double h = 0.86602540378443864676;
    const cplx w1 = make_double2(0.5, SIGN * h), w2 = make_double2(-0.5, SIGN * h);
    const cplx o1 = wf_mul(o[1], w1), o2 = wf_mul(o[2], w2);
    v[0] = wf_add(e[0], o[0]);
    v[3] = wf_sub(e[0], o[0]);
    v[1] = wf_add(e[1], o1);
    v[4] = wf_sub(e[1], o1);
    v[2] = wf_add(e[2], o2);
    v[5] = wf_sub(e[2], o2);
  } else if (R == 8) {
    cplx e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    wf_dft<4, SIGN>(e);
    wf_dft<4, SIGN>(o);
    const double r = 0.70710678118654752440;   // sqrt(2)/2
    // w8 = r (1 + SIGN i), w8^2 = SIGN i, w8^3 = r (-1 + SIGN i)
    const cplx o1 = make_double2(r * (o[1].x - SIGN * o[1].y), r * (o[1].y + SIGN * o[1].x));
    const cplx o2 = wf_si<SIGN>(o[2]);
    const cplx o3 = make_double2(r * (-o[3].x - SIGN * o[3].y), r * (-o[3].y + SIGN * o[3].x));
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

template <int R0, int SIGN>
struct WFFT {
  static constexpr int N = 64 * R0;
  static constexpr int NT = 8 * R0;          // active tasks
  static constexpr int LDS = 8 * R0 * WF_FP > R0 * WF_EP ? 8 * R0 * WF_FP : R0 * WF_EP;

  // ---- L -> T ----
  __host__ __device__ static void l2t_a(int l, cplx* a, cplx* E, const cplx* rt) {
    wf_dft<R0, SIGN>(a);
    E[l] = a[0];
#pragma unroll
    for (int k1 = 1; k1 < R0; ++k1) E[k1 * WF_EP + l] = wf_tw<SIGN>(a[k1], rt[l * k1]);
  }
  // task t' = l1 + 8 k1; reads E, the caller separates reads from writes
  __host__ __device__ static void l2t_b_read(int t, const cplx* E, cplx* d) {
    const int l1 = t & 7, k1 = t >> 3;
#pragma unroll
    for (int l2 = 0; l2 < 8; ++l2) d[l2] = E[k1 * WF_EP + l1 + 8 * l2];
  }
  __host__ __device__ static void l2t_b_write(int t, cplx* d, cplx* F, const cplx* rt) {
    const int l1 = t & 7, k1 = t >> 3;
    wf_dft<8, SIGN>(d);
    F[(0 * R0 + k1) * WF_FP + l1] = d[0];
#pragma unroll
    for (int q2 = 1; q2 < 8; ++q2)
      F[(q2 * R0 + k1) * WF_FP + l1] = wf_tw<SIGN>(d[q2], rt[l1 * q2 * R0]);
  }
  // task t = k1 + R0 q2: row t of F -> y[t + 8 R0 q1]
  __host__ __device__ static void l2t_c(int t, const cplx* F, cplx* out) {
#pragma unroll
    for (int l1 = 0; l1 < 8; ++l1) out[l1] = F[t * WF_FP + l1];
    wf_dft<8, SIGN>(out);
  }

  // ---- T -> L ----
  // task t = n1 + R0 q2 holds v[q1] = x[t + 8 R0 q1]; rows of F are
  // (l1, n1) with n1 fastest, which keeps these writes on distinct banks
  __host__ __device__ static void t2l_a(int t, cplx* v, cplx* F, const cplx* rt) {
    const int n1 = t % R0, q2 = t / R0;
    wf_dft<8, SIGN>(v);
    F[(0 * R0 + n1) * WF_FP + q2] = v[0];
#pragma unroll
    for (int l1 = 1; l1 < 8; ++l1)
      F[(l1 * R0 + n1) * WF_FP + q2] = wf_tw<SIGN>(v[l1], rt[q2 * l1 * R0]);
  }
  // task t' = l1 + 8 n1 reads row (l1, n1)
  __host__ __device__ static void t2l_b_read(int t, const cplx* F, cplx* u) {
    const int row = (t & 7) * R0 + (t >> 3);
#pragma unroll
    for (int q2 = 0; q2 < 8; ++q2) u[q2] = F[row * WF_FP + q2];
  }
  __host__ __device__ static void t2l_b_write(int t, cplx* u, cplx* E) {
    const int l1 = t & 7, n1 = t >> 3;
    wf_dft<8, SIGN>(u);
#pragma unroll
    for (int l2 = 0; l2 < 8; ++l2) E[n1 * WF_EP + l1 + 8 * l2] = u[l2];
  }
  // lane l -> a[m] = X[l + 64 m]
  __host__ __device__ static void t2l_c(int l, const cplx* E, cplx* a, const cplx* rt) {
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
template <int R0, int SIGN>
__device__ __forceinline__ void wfft_l2t(cplx* a, cplx* out, cplx* buf, const cplx* rt) {
  typedef WFFT<R0, SIGN> W;
  const int lane = threadIdx.x & 63;
  wf_sync();
  W::l2t_a(lane, a, buf, rt);
  wf_sync();
  cplx d[8];
  if (lane < W::NT) W::l2t_b_read(lane, buf, d);
  wf_sync();
  if (lane < W::NT) W::l2t_b_write(lane, d, buf, rt);
  wf_sync();
  if (lane < W::NT) W::l2t_c(lane, buf, out);
}

template <int R0, int SIGN>
__device__ __forceinline__ void wfft_t2l(cplx* v, cplx* a, cplx* buf, const cplx* rt) {
  typedef WFFT<R0, SIGN> W;
  const int lane = threadIdx.x & 63;
  wf_sync();
  if (lane < W::NT) W::t2l_a(lane, v, buf, rt);
  wf_sync();
  cplx u[8];
  if (lane < W::NT) W::t2l_b_read(lane, buf, u);
  wf_sync();
  if (lane < W::NT) W::t2l_b_write(lane, u, buf);
  wf_sync();
  W::t2l_c(lane, buf, a, rt);
}
