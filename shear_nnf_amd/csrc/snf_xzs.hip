// The x-row stage of hvhg (cross.f90:1415-1479, 2030-2045) for planes of
// Mx = 32 RX points (RX = 2, 3: 64 and 96, the ensemble boxes' sizes), on the
// half-wave FFT of snf_wfft32.h: a block of 4 waves takes 8 consecutive z-rows
// of one y, each 32-lane half of a wave one row.  The same work and the same
// results (to rounding) as the Stockham x-row kernel of snf_xz.hip, which it
// replaces for these plane sizes when it is the faster shape of a size class
// (its block shape is "threads 0, 8 rows" for the autotuner):
//   - the 6 hybrid fields of the block's rows staged through LDS (coalesced
//     runs of 8 z values, 128 bytes in FP64), three c2r x transforms per row;
//   - umax and the 'p' statistics (the 'p' partials in FP64, per block, as the
//     Stockham kernel writes them for blocks of 8 rows);
//   - u x omega, two r2c x transforms, H1/H2 separation, 3 hybrid rows out.
#include "snf_wfft32.h"

#define XZS_ROWS 8   // z-rows of a block: 4 waves x 2 halves

template <int RX, class H>
__device__ __forceinline__ void k_xzs_B_body(Geom g, const H* __restrict__ hin, H* __restrict__ hout,
                                             const cplx* __restrict__ twx, int rk,
                                             unsigned long long* umax_acc, double* __restrict__ phist,
                                             const DevScalars* sc, int nhist, int zblk, int j) {
  typedef WFFT32<RX, 1, H> WI;
  typedef typename wf_real<H>::T R;
  if (sc->halted) return;
  extern __shared__ double4 xzs_smem[];
  H* lds = reinterpret_cast<H*>(xzs_smem);
  // kxn = mgalx / 3 on these planes (a compile-time constant: the zero bins
  // between kxn and Mx - kxn drop out of the transforms)
  constexpr int Mx = 32 * RX, kxn = Mx / 3, SP = XZS_ROWS + 1;
  const int Nz = g.mgalz;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int half = lane >> 5, l = lane & 31;
  const int r = 2 * wave + half;          // this half's z-row in the block
  H* rt = lds;
  H* work = lds + Mx;                     // staging, then the halves' FFT buffers
  H* buf = work + r * WI::LDS;
  stage_roots(rt, twx, Mx);
  const int z0 = XZS_ROWS * zblk;
  const long fstride = (long)g.my * kxn * Nz;
  const long rowbase = (long)j * kxn * Nz + z0;   // + kx Nz + row
  // --- the block's tile: element i of a field is (kx, z) = (i / 8, i % 8) ---
  constexpr int TILE = kxn * XZS_ROWS;
  constexpr int NEF = (TILE + 255) / 256;
  const int i0 = threadIdx.x, kx0 = i0 >> 3, rr0 = i0 & 7;
  H pre[6][NEF];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const H* src = hin + f * fstride + rowbase + rr0;
#pragma unroll
    for (int q = 0; q < NEF; ++q)
      if (i0 + 256 * q < TILE) pre[f][q] = gld(src + (long)(kx0 + 32 * q) * Nz);
  }
  H* S = work;                            // [6][kxn][SP]
#pragma unroll
  for (int f = 0; f < 6; ++f)
#pragma unroll
    for (int q = 0; q < NEF; ++q)
      if (i0 + 256 * q < TILE) S[(f * kxn + kx0 + 32 * q) * SP + rr0] = pre[f][q];
  __syncthreads();
  // --- the three field pairs of row r in lane layout (bins l + 32 m) ---
  H zin[3][RX];
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const H* Sa = S + 2 * p * kxn * SP;
    const H* Sb = Sa + kxn * SP;
#pragma unroll
    for (int m = 0; m < RX; ++m) {
      const int b = l + 32 * m;
      H s = wf_mk((R)0, (R)0);
      if (32 * m >= kxn && 32 * m + 31 <= Mx - kxn) {
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
  __syncthreads();   // staging dead: the region becomes the FFT buffers
  // --- inverse x transforms: task t holds x = t + 4 RX q ---
  H zx[3][8];
#pragma unroll
  for (int p = 0; p < 3; ++p) wfft32_l2t<RX, 1>(zin[p], zx[p], buf, rt);
  const int ihist = ((sc->steps_local - 1) % nhist == 0) && rk == 1;
  const R c2 = (R)(2.0 * g.chi);
  const R sy = (R)(g.s * ((double)j * g.dy - 0.5 * g.Ly));
  R m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  double p0 = 0.0, p1 = 0.0, p2 = 0.0;
  H z12[8], z3[8];
  if (l < WI::NT) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const R u = zx[0][q].x, v = zx[0][q].y, w = zx[1][q].x, o1 = zx[1][q].y;
      const R o2 = zx[2][q].x, o3 = zx[2][q].y;
      z12[q] = wf_mk(v * (o3 + c2) - w * o2, w * o1 - u * (o3 + c2));
      z3[q] = wf_mk(u * o2 - v * o1, (R)0);
    }
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
  wfft32_t2l<RX, -1>(z12, a12, buf, rt);
  wfft32_t2l<RX, -1>(z3, a3, buf, rt);
  // separate H1, H2 from Z12 = H1 + i H2: bin Mx - k sits in another lane
  wf_sync();
#pragma unroll
  for (int m = 0; m < RX; ++m)
    if (32 * m < kxn || 32 * m + 31 > Mx - kxn) buf[l + 32 * m] = a12[m];
  wf_sync();
  __syncthreads();   // every half done with its FFT buffer before the output tile
  H* O = work;       // [3][kxn][SP]
  H h1[RX], h2[RX];
#pragma unroll
  for (int m = 0; m < RX; ++m) {
    const int k = l + 32 * m;
    if (k < kxn) {
      const H A = buf[k], Am = buf[k == 0 ? 0 : Mx - k];
      h1[m] = wf_mk((R)0.5 * (A.x + Am.x), (R)0.5 * (A.y - Am.y));
      h2[m] = wf_mk((R)0.5 * (A.y + Am.y), (R)-0.5 * (A.x - Am.x));
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < RX; ++m) {
    const int k = l + 32 * m;
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
      if (i0 + 256 * q < TILE) gst(dst + (long)(kx0 + 32 * q) * Nz, O[(f * kxn + kx0 + 32 * q) * SP + rr0]);
  }
  // --- reductions: block maxima into umax_acc, the 'p' sums of the block ---
  __shared__ double red[4][4];
  if (rk == 1) {
    for (int o = 32; o > 0; o >>= 1) {
      m0 = fmax(m0, __shfl_xor(m0, o, 64));
      m1 = fmax(m1, __shfl_xor(m1, o, 64));
      m2 = fmax(m2, __shfl_xor(m2, o, 64));
      m3 = fmax(m3, __shfl_xor(m3, o, 64));
    }
    if (lane == 0) {
      red[0][wave] = (double)m0;
      red[1][wave] = (double)m1;
      red[2][wave] = (double)m2;
      red[3][wave] = (double)m3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      double m = 0.0;
      for (int w = 0; w < 4; ++w) m = fmax(m, red[threadIdx.x][w]);
      // maxima of non-negative doubles order like their bit patterns
      atomicMax(&umax_acc[threadIdx.x], dbits(m));
    }
    __syncthreads();
  }
  if (ihist) {
    // the wave's two rows, then the block's 4 waves, in a fixed order
    for (int o = 32; o > 0; o >>= 1) {
      p0 += __shfl_xor(p0, o, 64);
      p1 += __shfl_xor(p1, o, 64);
      p2 += __shfl_xor(p2, o, 64);
    }
    if (lane == 0) {
      red[0][wave] = p0;
      red[1][wave] = p1;
      red[2][wave] = p2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      double acc = 0.0;
      for (int w = 0; w < 4; ++w) acc += red[threadIdx.x][w];
      phist[((long)j * (Nz / XZS_ROWS) + zblk) * 3 + threadIdx.x] = acc;
    }
  }
}

template <int RX, class H>
__global__ void __launch_bounds__(256)
k_xzs_B(Geom g, const H* __restrict__ hin, H* __restrict__ hout, const cplx* __restrict__ twx,
        int rk, unsigned long long* umax_acc, double* __restrict__ phist, const DevScalars* sc,
        int nhist) {
  const long lb = xcd_block((long)blockIdx.y * gridDim.x + blockIdx.x, (long)gridDim.x * gridDim.y);
  k_xzs_B_body<RX, H>(g, hin, hout, twx, rk, umax_acc, phist, sc, nhist, (int)(lb % gridDim.x),
                      (int)(lb / gridDim.x));
}

// LDS of k_xzs_B: the x roots and the larger of the 6-field staging tile and
// the 8 halves' FFT buffers (elements of es bytes)
inline size_t xzs_lds_B(int RX, size_t es) {
  const int Mx = 32 * RX, kxn = Mx / 3;
  const int R0 = RX, lw = R0 * WF32_EP > 4 * R0 * WF32_FP
                              ? (R0 * WF32_EP > 8 * R0 * WF32_GP ? R0 * WF32_EP : 8 * R0 * WF32_GP)
                              : (4 * R0 * WF32_FP > 8 * R0 * WF32_GP ? 4 * R0 * WF32_FP : 8 * R0 * WF32_GP);
  const size_t stage = (size_t)6 * kxn * (XZS_ROWS + 1), fft = (size_t)XZS_ROWS * lw;
  return (size_t)(Mx + (stage > fft ? stage : fft)) * es;
}
// the plane sizes k_xzs_B serves
inline bool xzs_size(int mgalx, int mgalz) {
  return (mgalx == 64 || mgalx == 96) && mgalz % XZS_ROWS == 0;
}

// ---------------------------------------------------------------------------
// The z-row stages for planes of Nz = 32 RZ points (RZ = 2, 3), as the wave
// kernels A and C of snf_xzw.hip with a (kx, y) row per 32-lane half: a block
// of 4 waves takes 8 consecutive y of one kx (the kz chunks' runs of 8 x 4
// values), blocks in order of y then kx.  A half whose y lies beyond my
// computes on the last row and stores nothing.
__device__ __forceinline__ void xzs_row(const Geom& g, int blk, int rr, int& j, int& i, bool& on) {
  const int nyg = (g.my + XZS_ROWS - 1) / XZS_ROWS;
  i = blk / nyg;
  j = (blk - i * nyg) * XZS_ROWS + rr;
  on = j < g.my;
  if (!on) j = g.my - 1;
}
inline unsigned xzs_rows_blocks(const Geom& g) {
  return (unsigned)(g.kxn * ((g.my + XZS_ROWS - 1) / XZS_ROWS));
}

template <int RZ, class H>
__device__ __forceinline__ void k_xzs_A_body(Geom g, const cplx* __restrict__ vor, const cplx* __restrict__ phi,
       const cplx* __restrict__ hgf, const cplx* __restrict__ hvf, const cplx* __restrict__ dvordy,
       const double* __restrict__ u00, const double* __restrict__ w00, const double* __restrict__ dudy,
       const double* __restrict__ dwdy, double vbulk, H* __restrict__ hyb,
       const cplx* __restrict__ twz, const DevScalars* sc, int blk) {
  typedef WFFT32<RZ, 1, H> W;
  typedef typename wf_real<H>::T R;
  if (sc->halted) return;
  extern __shared__ double4 xzs_smem[];
  H* lds = reinterpret_cast<H*>(xzs_smem);
  constexpr int Nz = 32 * RZ, NZ = wz_nz(Nz), MZ = wz_mz(Nz);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, l = lane & 31;
  const int rr = 2 * wave + (lane >> 5);
  H* rt = lds;
  H* buf = lds + Nz + rr * W::LDS;
  stage_roots(rt, twz, Nz);
  __syncthreads();
  int j, i;
  bool on;
  xzs_row(g, blk, rr, j, i, on);
  const double u00j = gldd(u00 + j), w00j = gldd(w00 + j);
  const double dudyj = gldd(dudy + j), dwdyj = gldd(dwdy + j);
  H d[6][RZ];
#pragma unroll
  for (int m = 0; m < RZ; ++m) {
    const int k = wz_k(l + 32 * m, NZ, MZ, Nz);
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
      q.v = ld_io(hgf, idx, g.io32 & 1);
      q.dv = ld_io(hvf, idx, g.io32 & 1);
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
  H* dst = hyb + ((long)j * g.kxn + i) * Nz;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    H out[8];
    wfft32_l2t<RZ, 1>(d[f], out, buf, rt);
    if (on && l < W::NT) {
#pragma unroll
      for (int q = 0; q < 8; ++q) gst(dst + f * fstride + l + 4 * RZ * q, out[q]);
    }
  }
}

template <int RZ, class H>
__device__ __forceinline__ void k_xzs_C_body(Geom g, const H* __restrict__ hout, const cplx* __restrict__ vel,
       cplx* __restrict__ dvordy, cplx* __restrict__ hgf, cplx* __restrict__ hvf,
       double* __restrict__ rf0u, double* __restrict__ rf0w, const cplx* __restrict__ twz,
       const DevScalars* sc, int blk) {
  typedef WFFT32<RZ, -1, H> W;
  if (sc->halted) return;
  extern __shared__ double4 xzs_smem[];
  H* lds = reinterpret_cast<H*>(xzs_smem);
  constexpr int Nz = 32 * RZ, NZ = wz_nz(Nz), MZ = wz_mz(Nz);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, l = lane & 31;
  const int rr = 2 * wave + (lane >> 5);
  H* rt = lds;
  H* buf = lds + Nz + rr * W::LDS;
  stage_roots(rt, twz, Nz);
  __syncthreads();
  int j, i;
  bool on;
  xzs_row(g, blk, rr, j, i, on);
  const long fstride = (long)g.my * g.kxn * Nz;
  const H* src = hout + ((long)j * g.kxn + i) * Nz;
  H h[3][RZ];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    H v[8];
    if (l < W::NT) {
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = gld(src + f * fstride + l + 4 * RZ * q);
    }
    wfft32_t2l<RZ, -1>(v, h[f], buf, rt);
  }
  if (!on) return;
  const double norm = 1.0 / ((double)g.mgalx * (double)Nz);
  const double kx = g.alp * (double)i;
  const double kx2 = kx * kx;
#pragma unroll
  for (int m = 0; m < RZ; ++m) {
    const int k = wz_k(l + 32 * m, NZ, MZ, Nz);
    if (k >= 0) {
      const cplx h1 = crmul(norm, to_d(h[0][m])), h2 = crmul(norm, to_d(h[1][m])),
                 h3 = crmul(norm, to_d(h[2][m]));
      const double kz = kz_of(k, NZ, MZ, g.gam);
      const long idx = sidx(g, j, i, k);
      const cplx v = ld_io(vel, idx, g.io32 & 1);
      st_io(dvordy, idx, cadd(mul_i(kx, h1), mul_i(kz, h3)), g.io32 & 2);
      const cplx dd = mkc(h1.x - g.s * v.x, h1.y - g.s * v.y);
      st_io(hgf, idx, cadd(mul_i(-kx, h3), mul_i(kz, dd)), g.io32 & 2);
      const double kk = kx2 + kz * kz;
      st_io(hvf, idx, mkc(-h2.x * kk, -h2.y * kk), g.io32 & 2);
      if (i == 0 && k == 0) {
        gstd(rf0u + j, h1.x);
        gstd(rf0w + j, h3.x);
      }
    }
  }
}

template <int RZ, class H>
__global__ void __launch_bounds__(256)
k_xzs_A(Geom g, const cplx* __restrict__ vor, const cplx* __restrict__ phi,
        const cplx* __restrict__ hgf, const cplx* __restrict__ hvf, const cplx* __restrict__ dvordy,
        const double* __restrict__ u00, const double* __restrict__ w00, const double* __restrict__ dudy,
        const double* __restrict__ dwdy, double vbulk, H* __restrict__ hyb,
        const cplx* __restrict__ twz, const DevScalars* sc) {
  k_xzs_A_body<RZ, H>(g, vor, phi, hgf, hvf, dvordy, u00, w00, dudy, dwdy, vbulk, hyb, twz, sc,
                      (int)xcd_block(blockIdx.x, gridDim.x));
}
template <int RZ, class H>
__global__ void __launch_bounds__(256)
k_xzs_C(Geom g, const H* __restrict__ hout, const cplx* __restrict__ vel, cplx* __restrict__ dvordy,
        cplx* __restrict__ hgf, cplx* __restrict__ hvf, double* __restrict__ rf0u,
        double* __restrict__ rf0w, const cplx* __restrict__ twz, const DevScalars* sc) {
  k_xzs_C_body<RZ, H>(g, hout, vel, dvordy, hgf, hvf, rf0u, rf0w, twz, sc,
                      (int)xcd_block(blockIdx.x, gridDim.x));
}
// LDS of k_xzs_A and k_xzs_C: the z roots and the 8 halves' FFT buffers
inline size_t xzs_lds_AC(int RZ, size_t es) {
  const int R0 = RZ, lw = R0 * WF32_EP > 4 * R0 * WF32_FP
                              ? (R0 * WF32_EP > 8 * R0 * WF32_GP ? R0 * WF32_EP : 8 * R0 * WF32_GP)
                              : (4 * R0 * WF32_FP > 8 * R0 * WF32_GP ? 4 * R0 * WF32_FP : 8 * R0 * WF32_GP);
  return (size_t)(32 * RZ + XZS_ROWS * lw) * es;
}
// the z-plane sizes k_xzs_A and k_xzs_C serve
inline bool xzs_z_size(int mgalz) { return mgalz == 64 || mgalz == 96; }
