// Nonlinear stage of one RK3 substep (hvhg, cross.f90:1127-1765) for planes
// of 64*R points in x and z (R = 2, 3, 4, 6, 8), built on the wave-level FFT of
// snf_wfft.h: every transform runs inside one wave, in registers, with two
// wave-private LDS exchanges and no block barrier.
//
//   WA  one wave per (kx, y) row, a block 4 consecutive y of one kx: loads
//       the 5 spectral inputs with the fill00 map, KMM reconstruction of u,
//       v, w, omega (cross.f90:1221-1276), inverse z transform of the 6
//       fields -> hybrid [f][y][kx][z]
//   WB  one block per (4 z-rows, y), one wave per row: the 6 hybrid fields
//       staged through LDS (coalesced 64-byte runs of 4 z values), three
//       paired c2r x transforms, u x omega, umax and the 'p' statistics
//       (cross.f90:1430-1462, 2030-2045), r2c x transforms -> hybrid, 3 fields
//   WC  one wave per (kx, y) row, blocks as WA: forward z transform,
//       pack00, normalisation 1/(mgalx mgalz), output combinations and
//       rf0u/rf0w (cross.f90:1477-1546)
// Same data layout and the same results (to rounding) as the generic
// Stockham kernels of snf_xz.hip, which serve every other grid size.
#include "snf_internal.h"
#include "snf_wfft.h"

#define WA_WAVES 4
#ifndef WB_ROWS
#define WB_ROWS 4
#endif
// z-rows of an x-row block: 4 in FP64 (the 4-row tile's 64-byte runs), 8 in
// FP32 (64-byte runs again; 4 rows of 32 bytes took 0.437 ms against 0.430 at
// 384x385x384)
template <class H> __host__ __device__ constexpr int wb_rows() { return sizeof(H) == 8 ? 8 : WB_ROWS; }
#define WC_WAVES 4
#define UMAX_SLOTS 64   // umax accumulators of the x-row kernel, folded by k_scalar

// kz modes of a plane of Nz = mgalz points: mz = 2 (Nz / 3) - 1, nz = (mz - 1) / 2
// (compile-time constants in the kernels of this file, where Nz = 64 R)
__host__ __device__ constexpr int wz_mz(int Nz) { return 2 * (Nz / 3) - 1; }
__host__ __device__ constexpr int wz_nz(int Nz) { return Nz / 3 - 1; }
// padded z position (0..Nz-1) of lane l, value m -> kz index, or -1 (gap)
__device__ __forceinline__ int wz_k(int pos, int nz, int mz, int Nz) {
  if (pos <= nz) return pos;
  if (pos >= Nz - nz) return pos - Nz + mz;
  return -1;
}

// Row (j, i) of wave w in the row kernels A and C: a block takes WAVES
// consecutive y of one kx, so that together its waves read whole runs of
// the kz chunks (WAVES * SNF_CW elements of each); consecutive blocks, on
// one XCD, take the next y.  Grid: kxn * ceil(my / WAVES) blocks.
__device__ __forceinline__ bool wrow(const Geom& g, int waves, int w, int& j, int& i) {
  const long nyg = (g.my + waves - 1) / waves;
  const long lb = xcd_block(blockIdx.x, gridDim.x);
  i = (int)(lb / nyg);
  j = (int)(lb - (long)i * nyg) * waves + w;
  return i < g.kxn && j < g.my;
}
inline unsigned wrow_blocks(const Geom& g, int waves) {
  return (unsigned)((long)g.kxn * ((g.my + waves - 1) / waves));
}

template <class H>
__device__ __forceinline__ void stage_roots(H* rt, const cplx* tw, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) rt[t] = to_h<H>(gld(tw + t));
}

// ---------------------------------------------------------------------------
template <int RZ, class H>
__global__ void __launch_bounds__(64 * WA_WAVES)
k_xzw_A(Geom g, const cplx* __restrict__ vor, const cplx* __restrict__ phi,
        const cplx* __restrict__ hg, const cplx* __restrict__ hv,
        const cplx* __restrict__ dvordy, const double* __restrict__ u00,
        const double* __restrict__ w00, const double* __restrict__ dudy,
        const double* __restrict__ dwdy, double vbulk, H* __restrict__ hyb,
        const cplx* __restrict__ twz, const DevScalars* sc) {
  typedef WFFT<RZ, 1, H> W;
  typedef typename wf_real<H>::T R;
  if (sc->halted) return;
  extern __shared__ double4 xzw_smem[];
  H* lds = reinterpret_cast<H*>(xzw_smem);
  H* rt = lds;
  constexpr int Nz = 64 * RZ, NZ = wz_nz(Nz), MZ = wz_mz(Nz);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  H* buf = lds + Nz + wave * W::LDS;
  stage_roots(rt, twz, Nz);
  __syncthreads();
  int j, i;
  if (!wrow(g, WA_WAVES, wave, j, i)) return;
  const long row = (long)j * g.kxn + i;
  const double u00j = gldd(u00 + j), w00j = gldd(w00 + j);
  const double dudyj = gldd(dudy + j), dwdyj = gldd(dwdy + j);
  H d[6][RZ];
#pragma unroll
  for (int m = 0; m < RZ; ++m) {
    const int k = wz_k(lane + 64 * m, NZ, MZ, Nz);
    if (k >= 0) {
      const long idx = sidx(g, j, i, k);
      PlaneIn q;
      if (i == 0 && k > NZ) {     // set_conj (operators.f90:1382-1429)
        const long cj = sidx(g, j, 0, MZ - k);
        q.phic = cconj(gld(phi + cj));
        q.ome2c = cconj(gld(vor + cj));
      } else {
        q.phic = gld(phi + idx);
        q.ome2c = gld(vor + idx);
      }
      q.v = ld_io(hg, idx, g.io32 & 1);
      q.dv = ld_io(hv, idx, g.io32 & 1);
      q.dome = ld_io(dvordy, idx, g.io32 & 1);
      cplx o[6];
      derived6(g, q, i, k, u00j, w00j, dudyj, dwdyj, vbulk, o);
#pragma unroll
      for (int f = 0; f < 6; ++f) d[f][m] = to_h<H>(o[f]);
    } else {
#pragma unroll
      for (int f = 0; f < 6; ++f) d[f][m] = wf_mk((R)0, (R)0);
    }
  }
  const long fstride = (long)g.my * g.kxn * Nz;
  H* dst = hyb + row * Nz;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    H out[8];
    wfft_l2t<RZ, 1>(d[f], out, buf, rt);
    if (lane < W::NT) {
#pragma unroll
      for (int q = 0; q < 8; ++q) gst(dst + f * fstride + lane + 8 * RZ * q, out[q]);
    }
  }
}

// ---------------------------------------------------------------------------
template <int RX, class H, int ROWS = wb_rows<H>()>
__global__ void __launch_bounds__(64 * ROWS)
k_xzw_B(Geom g, const H* __restrict__ hin, H* __restrict__ hout,
        const cplx* __restrict__ twx, int rk, double* __restrict__ upart,
        double* __restrict__ phist, const DevScalars* sc, int nhist) {
  typedef WFFT<RX, 1, H> WI;
  typedef typename wf_real<H>::T R;
  if (sc->halted) return;
  extern __shared__ double4 xzw_smem[];
  H* lds = reinterpret_cast<H*>(xzw_smem);
  // kxn = mgalx / 3 on every plane of this path: a compile-time constant, so
  // that the dealiasing gaps (bins kxn .. Mx - kxn, zero in and unused out)
  // drop out of the transforms' first and last stages
  constexpr int Mx = 64 * RX, kxn = Mx / 3;
  const int Nz = g.mgalz;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // planes up to 384 points (WIDE) stage all 6 fields in one round and run
  // independent transforms in pairs; the LDS this takes (67.5 KB at 384,
  // FP64) still leaves 2 blocks per CU, which 512-point planes would lose
  constexpr bool WIDE = RX <= 6;
  constexpr int NFR = WIDE ? 6 : 2;      // fields staged per round
  H* rt = lds;
  H* work = lds + Mx;                    // staging, then the FFT buffers of each wave
  H* buf = work + wave * (WIDE ? 2 : 1) * WI::LDS;
  H* buf1 = buf + WI::LDS;               // second buffer (WIDE)
  stage_roots(rt, twx, Mx);
  const long lb = xcd_block((long)blockIdx.y * gridDim.x + blockIdx.x, (long)gridDim.x * gridDim.y);
  const int zblk = (int)(lb % gridDim.x), j = (int)(lb / gridDim.x);
  const int z0 = ROWS * zblk, r = wave;
  const long fstride = (long)g.my * kxn * Nz;
  const long rowbase = (long)j * kxn * Nz + z0;   // + kx*Nz + r
  // --- load the 3 field pairs of row r in lane layout (bins l + 64 m) ---
  // every load of the block is issued up front (runs of 4 z values, 64 B),
  // field by field: element i of a field is (kx, z) = (i / ROWS, i % ROWS);
  // the tile then passes through LDS, NFR fields at a time
  constexpr int KXM = kxn;
  constexpr int NEF = (KXM * ROWS + 64 * ROWS - 1) / (64 * ROWS);   // per field
  const int tile = kxn * ROWS;                                   // per field
  const int i0 = threadIdx.x, kx0 = i0 / ROWS, rr0 = i0 % ROWS;
  constexpr int KXS = 64;                                           // kx step per q
  H pre[6][NEF];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const H* src = hin + f * fstride + rowbase + rr0;
#pragma unroll
    for (int q = 0; q < NEF; ++q)
      if (i0 + q * 64 * ROWS < tile) pre[f][q] = gld(src + (long)(kx0 + q * KXS) * Nz);
  }
  H zin[3][RX];
  H* S = work;                           // [NFR][kxn][ROWS + 1]
  const int SP = ROWS + 1;
#pragma unroll
  for (int f0 = 0; f0 < 6; f0 += NFR) {
    __syncthreads();
#pragma unroll
    for (int h = 0; h < NFR; ++h)
#pragma unroll
      for (int q = 0; q < NEF; ++q)
        if (i0 + q * 64 * ROWS < tile) S[(h * kxn + kx0 + q * KXS) * SP + rr0] = pre[f0 + h][q];
    __syncthreads();
#pragma unroll
    for (int pp = 0; pp < NFR / 2; ++pp) {
      const int p = f0 / 2 + pp;
      const H* Sa = S + 2 * pp * kxn * SP;
      const H* Sb = Sa + kxn * SP;
#pragma unroll
      for (int m = 0; m < RX; ++m) {
        const int b = lane + 64 * m;
        H s = wf_mk((R)0, (R)0);
        if (64 * m >= kxn && 64 * m + 63 <= Mx - kxn) {
          // a row of the dealiasing gap: zero
        } else if (b == 0) {
          const H Ga = Sa[r], Gb = Sb[r];
          s = wf_mk(Ga.x, Gb.x);   // c2r ignores the kx=0 imaginary part
        } else if (b < kxn) {
          const H Ga = Sa[b * SP + r], Gb = Sb[b * SP + r];
          s = wf_mk(Ga.x - Gb.y, Ga.y + Gb.x);
        } else if (b > Mx - kxn) {
          const int i = Mx - b;
          const H Ga = Sa[i * SP + r], Gb = Sb[i * SP + r];
          s = wf_mk(Ga.x + Gb.y, Gb.x - Ga.y);
        }
        zin[p][m] = s;
      }
    }
  }
  __syncthreads();   // staging dead: the region becomes the FFT buffers
  // --- inverse x transforms: task t holds x = t + 8 RX q ---
  H zx[3][8];
  if (WIDE) {
    wfft_l2t2<RX, 1>(zin[0], zin[1], zx[0], zx[1], buf, buf1, rt);
  } else {
    wfft_l2t<RX, 1>(zin[0], zx[0], buf, rt);
    wfft_l2t<RX, 1>(zin[1], zx[1], buf, rt);
  }
  wfft_l2t<RX, 1>(zin[2], zx[2], buf, rt);
  const int ihist = ((sc->steps_local - 1) % nhist == 0) && rk == 1;
  const R c2 = (R)(2.0 * g.chi);
  const R sy = (R)(g.s * ((double)j * g.dy - 0.5 * g.Ly));
  // maxima in the element type, the 'p' sums in FP64
  R m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  double p0 = 0.0, p1 = 0.0, p2 = 0.0;
  H z12[8], z3[8];
  if (lane < WI::NT) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const R u = zx[0][q].x, v = zx[0][q].y, w = zx[1][q].x, o1 = zx[1][q].y;
      const R o2 = zx[2][q].x, o3 = zx[2][q].y;
      z12[q] = wf_mk(v * (o3 + c2) - w * o2, w * o1 - u * (o3 + c2));
      z3[q] = wf_mk(u * o2 - v * o1, (R)0);
    }
    // the maxima only at rk 1, the 'p' sums only on a statistics step
    // (uniform branches: the other substeps skip the work)
    if (rk == 1) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const R u = zx[0][q].x, v = zx[0][q].y, w = zx[1][q].x;
        m0 = fmax(m0, fabs(u));
        m1 = fmax(m1, fabs(u + sy));
        m2 = fmax(m2, fabs(v));
        m3 = fmax(m3, fabs(w));
      }
    }
    if (ihist) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const R u = zx[0][q].x, v = zx[0][q].y, w = zx[1][q].x;
        p0 += (double)(v * (u * u));
        p1 += (double)(v * (w * w));
        p2 += (double)(v * v * v);
      }
    }
  }
  // --- forward x transforms -> lane layout bins ---
  H a12[RX], a3[RX];
  if (WIDE) {
    wfft_t2l2<RX, -1>(z12, z3, a12, a3, buf, buf1, rt);
  } else {
    wfft_t2l<RX, -1>(z12, a12, buf, rt);
    wfft_t2l<RX, -1>(z3, a3, buf, rt);
  }
  // separate H1, H2 from Z12 = H1 + i H2: needs bin Mx - k from another lane
  // (only the rows of bins below kxn or above Mx - kxn are read)
  wf_sync();
#pragma unroll
  for (int m = 0; m < RX; ++m)
    if (64 * m < kxn || 64 * m + 63 > Mx - kxn) buf[lane + 64 * m] = a12[m];
  wf_sync();
  __syncthreads();   // every wave done with its FFT buffer before the output tile
  H* O = work;       // [3][kxn][ROWS + 1]
  H h1[RX], h2[RX];
#pragma unroll
  for (int m = 0; m < RX; ++m) {
    const int k = lane + 64 * m;
    if (k < kxn) {
      const H A = buf[k], Am = buf[k == 0 ? 0 : Mx - k];
      h1[m] = wf_mk((R)0.5 * (A.x + Am.x), (R)0.5 * (A.y - Am.y));
      h2[m] = wf_mk((R)0.5 * (A.y + Am.y), (R)-0.5 * (A.x - Am.x));
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < RX; ++m) {
    const int k = lane + 64 * m;
    if (k < kxn) {
      O[(0 * kxn + k) * SP + r] = h1[m];
      O[(1 * kxn + k) * SP + r] = h2[m];
      O[(2 * kxn + k) * SP + r] = a3[m];
    }
  }
  __syncthreads();
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    H* dst = hout + f * fstride + rowbase + rr0;
#pragma unroll
    for (int q = 0; q < NEF; ++q)
      if (i0 + q * 64 * ROWS < tile)
        gst(dst + (long)(kx0 + q * KXS) * Nz, O[(f * kxn + kx0 + q * KXS) * SP + rr0]);
  }
  // --- reductions: one atomic per wave for umax, one partial per row ---
  if (rk == 1) {
    for (int o = 32; o > 0; o >>= 1) {
      m0 = fmax(m0, __shfl_xor(m0, o, 64));
      m1 = fmax(m1, __shfl_xor(m1, o, 64));
      m2 = fmax(m2, __shfl_xor(m2, o, 64));
      m3 = fmax(m3, __shfl_xor(m3, o, 64));
    }
    __shared__ double um[ROWS][4];
    if (lane == 0) {
      um[wave][0] = (double)m0;
      um[wave][1] = (double)m1;
      um[wave][2] = (double)m2;
      um[wave][3] = (double)m3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {   // one partial per block
      double m = 0.0;
#pragma unroll
      for (int w = 0; w < ROWS; ++w) m = fmax(m, um[w][threadIdx.x]);
      // maxima of non-negative doubles order like their bit patterns
      atomicMax(reinterpret_cast<unsigned long long*>(upart) + (lb % UMAX_SLOTS) * 4 + threadIdx.x,
                (unsigned long long)__double_as_longlong(m));
    }
  }
  if (ihist) {
    for (int o = 32; o > 0; o >>= 1) {
      p0 += __shfl_xor(p0, o, 64);
      p1 += __shfl_xor(p1, o, 64);
      p2 += __shfl_xor(p2, o, 64);
    }
    if (lane == 0) {
      double* ph = phist + ((long)j * Nz + z0 + r) * 3;
      gstd(ph, p0);
      gstd(ph + 1, p1);
      gstd(ph + 2, p2);
    }
  }
}

// ---------------------------------------------------------------------------
template <int RZ, class H>
__global__ void __launch_bounds__(64 * WC_WAVES)
k_xzw_C(Geom g, const H* __restrict__ hout, const cplx* __restrict__ vel,
        cplx* __restrict__ dvordy, cplx* __restrict__ hg, cplx* __restrict__ hv,
        double* __restrict__ rf0u, double* __restrict__ rf0w, const cplx* __restrict__ twz,
        const DevScalars* sc) {
  typedef WFFT<RZ, -1, H> W;
  if (sc->halted) return;
  extern __shared__ double4 xzw_smem[];
  H* lds = reinterpret_cast<H*>(xzw_smem);
  H* rt = lds;
  constexpr int Nz = 64 * RZ, NZ = wz_nz(Nz), MZ = wz_mz(Nz);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  H* buf = lds + Nz + wave * W::LDS;
  stage_roots(rt, twz, Nz);
  __syncthreads();
  int j, i;
  if (!wrow(g, WC_WAVES, wave, j, i)) return;
  const long row = (long)j * g.kxn + i;
  const long fstride = (long)g.my * g.kxn * Nz;
  const H* src = hout + row * Nz;
  H h[3][RZ];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    H v[8];
    if (lane < W::NT) {
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = gld(src + f * fstride + lane + 8 * RZ * q);
    }
    wfft_t2l<RZ, -1>(v, h[f], buf, rt);
  }
  const double norm = 1.0 / ((double)g.mgalx * (double)Nz);
  const double kx = g.alp * (double)i;
  const double kx2 = kx * kx;
#pragma unroll
  for (int m = 0; m < RZ; ++m) {
    const int k = wz_k(lane + 64 * m, NZ, MZ, Nz);
    if (k >= 0) {
      const cplx h1 = crmul(norm, to_d(h[0][m])), h2 = crmul(norm, to_d(h[1][m])),
                 h3 = crmul(norm, to_d(h[2][m]));
      const double kz = kz_of(k, NZ, MZ, g.gam);
      const long idx = sidx(g, j, i, k);
      const cplx v = ld_io(vel, idx, g.io32 & 1);
      st_io(dvordy, idx, cadd(mul_i(kx, h1), mul_i(kz, h3)), g.io32 & 2);
      const cplx dd = mkc(h1.x - g.s * v.x, h1.y - g.s * v.y);
      st_io(hg, idx, cadd(mul_i(-kx, h3), mul_i(kz, dd)), g.io32 & 2);
      const double kk = kx2 + kz * kz;
      st_io(hv, idx, mkc(-h2.x * kk, -h2.y * kk), g.io32 & 2);
      if (i == 0 && k == 0) {
        gstd(rf0u + j, h1.x);
        gstd(rf0w + j, h3.x);
      }
    }
  }
}

// LDS bytes of the three kernels for R = Nz/64 (A, C) and Rx = Mx/64 (B),
// elements of es bytes (16: FP64, 8: FP32)
inline size_t xzw_lds_A(int R, size_t es) { return (size_t)(64 * R + WA_WAVES * (R * WF_EP > 8 * R * WF_FP ? R * WF_EP : 8 * R * WF_FP)) * es; }
inline size_t xzw_lds_C(int R, size_t es) { return xzw_lds_A(R, es); }
inline size_t xzw_lds_B(int Rx, int kxn, size_t es, int rows = WB_ROWS) {
  const bool wide = Rx <= 6;   // k_xzw_B's WIDE
  const size_t fft = (size_t)rows * (wide ? 2 : 1) * (Rx * WF_EP > 8 * Rx * WF_FP ? Rx * WF_EP : 8 * Rx * WF_FP);
  const size_t stage = (size_t)(wide ? 6 : 2) * kxn * (rows + 1);
  const size_t out = (size_t)3 * kxn * (rows + 1);
  size_t w = fft > stage ? fft : stage;
  w = w > out ? w : out;
  return (size_t)(64 * Rx + w) * es;
}
