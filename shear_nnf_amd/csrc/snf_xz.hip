// Nonlinear stage of one RK3 substep (hvhg, cross.f90:1127-1765):
// velocity/vorticity reconstruction, inverse x-z transforms, u x omega,
// forward transforms and the three output combinations.
//
// Data movement per substep (Nz = mgalz, Mx = mgalx):
//   A  spectral (5 inputs, kz-chunk layout) -> hybrid [f][y][kx][z], 6 fields,
//      z-inverse DFT per (kx, y) row      (grid my / rows x kxn)
//   B  hybrid rows (all kx at one z)   -> physical x rows in LDS,
//      u x omega, x-forward DFT -> hybrid [f][y][kx][z], 3 fields
//                                                  (grid Nz/2 x my)
//   C  hybrid columns -> z-forward DFT -> 3 spectral outputs
//                                         (grid my / rows x kxn)
// The kx=0 real-to-complex convention of a c2r transform (imaginary part
// of the kx=0 coefficient ignored) and the 3/2-rule mode map of
// fill00/pack00 (fou3D_thread.f:75-161) are reproduced exactly.
#include "snf_internal.h"
#include "snf_fft.h"

__device__ __forceinline__ int kz_pos(int k, int nz, int mz, int Nz) {
  return k <= nz ? k : k - mz + Nz;
}
// (0 + i c) * a
__device__ __forceinline__ cplx mul_i(double c, cplx a) { return mkc(-c * a.y, c * a.x); }

// Spectral inputs of plane j at (i,k) with the kx=0 Hermitian fix of
// set_conj applied to phi and omega_y (operators.f90:1382-1429).
struct PlaneIn {
  cplx phic, ome2c, v, dv, dome;
};

__device__ __forceinline__ PlaneIn load_in(const Geom& g, const cplx* vor, const cplx* phi,
                                           const cplx* hg, const cplx* hv, const cplx* dvordy,
                                           int j, int i, int k) {
  PlaneIn r;
  const long idx = sidx(g, j, i, k);
  if (i == 0 && k > g.nz) {
    const long cj = sidx(g, j, 0, g.mz - k);
    r.phic = cconj(phi[cj]);
    r.ome2c = cconj(vor[cj]);
  } else {
    r.phic = phi[idx];
    r.ome2c = vor[idx];
  }
  r.v = ld_io(hg, idx, g.io32 & 1);
  r.dv = ld_io(hv, idx, g.io32 & 1);
  r.dome = ld_io(dvordy, idx, g.io32 & 1);
  return r;
}

// u, v, w, omega_x, omega_y, omega_z at (i,k) (cross.f90:1221-1276).
__device__ __forceinline__ void derived6(const Geom& g, const PlaneIn& q, int i, int k,
                                         double u00j, double w00j, double dudyj,
                                         double dwdyj, double vbulk, cplx* out) {
  const double kx = g.alp * (double)i;
  const double kz = kz_of(k, g.nz, g.mz, g.gam);
  cplx u, w, o1, o3;
  if (i == 0 && k == 0) {
    u = mkc(u00j, 0.0);
    w = mkc(w00j, 0.0);
    o3 = mkc(-dudyj, 0.0);
    o1 = mkc(dwdyj, 0.0);
  } else if (i == 0) {
    // one correctly rounded reciprocal, then products (within 1.5 ulp of the
    // quotients; 8 IEEE divisions per value took a fifth of A's FP64 work)
    // (x / (i kz) = i (-1/kz) x)
    const double rz = 1.0 / kz;
    o3 = mul_i(rz, q.dome);      // -dome / (i kz)
    o1 = mul_i(rz, q.phic);      // -phic / (i kz)
    w = mul_i(rz, q.dv);         // -dv / (i kz)
    u = mul_i(-rz, q.ome2c);     // ome2c / (i kz)
  } else {
    const double rk = 1.0 / (kx * kx + kz * kz);
    const cplx a1 = mul_i(kz, q.dome), a2 = mul_i(kx, q.phic);
    o3 = mkc((a1.x - a2.x) * rk, (a1.y - a2.y) * rk);
    const cplx b1 = mul_i(kx, q.dome), b2 = mul_i(kz, q.phic);
    o1 = mkc((b1.x + b2.x) * rk, (b1.y + b2.y) * rk);
    const cplx c1 = mul_i(kz, q.dv), c2 = mul_i(kx, q.ome2c);
    w = mkc((c1.x + c2.x) * rk, (c1.y + c2.y) * rk);
    const cplx d1 = mul_i(kx, q.dv), d2 = mul_i(kz, q.ome2c);
    u = mkc((d1.x - d2.x) * rk, (d1.y - d2.y) * rk);
  }
  out[0] = u;
  out[1] = (i == 0 && k == 0) ? mkc(vbulk, 0.0) : q.v;
  out[2] = w;
  out[3] = o1;
  out[4] = q.ome2c;
  out[5] = o3;
}

// ---------------------------------------------------------------------------
// A: spectral -> hybrid (inverse z DFT) of u,v,w,o1,o2,o3; xz_rows_per_block
// (kx, y) rows per block (grid my / rows x kxn), two groups of three sequences
// per row (ping-pong LDS buffers)
// ---------------------------------------------------------------------------
#define XZ_MAXV 12       // values per thread per FFT pass in the x-row kernel
#define XZ_B_MAXW 2      // (row pair, x) product items per thread


// Rows (kx, y) per block of the Stockham A and C kernels: enough that the
// block's 3 * rows sequences of Nz points keep its 256 threads busy (short
// planes), at most 8.  The host sizes the grid (my / rows blocks per kx)
// and the LDS (2 * 3 * rows * Nz elements and Nz FP64 roots) with the same
// function.
// t / d for 0 <= t < 2^16 and 1 <= d <= 1024 from rd = 1/d in FP32: (t + 1/2) / d
// lies at least 1/(2d) from an integer, and the two roundings move it by at
// most (t/d) 2^-22, so the truncation is exact (no integer division)
__device__ __forceinline__ int xz_div(int t, float rd) { return (int)(((float)t + 0.5f) * rd); }

__host__ __device__ __forceinline__ int xz_rows_per_block(int Nz) {
  const int r = 1536 / (3 * Nz);
  return r < 1 ? 1 : (r > 8 ? 8 : r);
}

// H: element type of the transforms and hybrid planes (cplx, or cplxf in the
// mixed precision mode); the spectral fields and the reconstruction are FP64.
// A block takes rows y = j0 .. j0 + rows - 1 of one kx; its loads run over
// (kz chunk, row, kz % SNF_CW), so consecutive threads read consecutive
// elements of the kz-chunk layout.
template <class H>
__device__ __forceinline__ void k_xz_A_body(Geom g, const cplx* __restrict__ vor, const cplx* __restrict__ phi,
       const cplx* __restrict__ hg, const cplx* __restrict__ hv,
       const cplx* __restrict__ dvordy, const double* __restrict__ u00,
       const double* __restrict__ w00, const double* __restrict__ dudy,
       const double* __restrict__ dwdy, double vbulk, H* __restrict__ hyb, FFTPlan pz,
       const cplx* __restrict__ twz, const DevScalars* sc) {
  typedef typename wf_real<H>::T R;
  if (sc->halted) return;
  extern __shared__ double4 xz_smem[];
  const int Nz = g.mgalz, RB = xz_rows_per_block(Nz);
  H* a = reinterpret_cast<H*>(xz_smem);
  H* b = a + 3 * RB * Nz;
  cplx* tw = reinterpret_cast<cplx*>(b + 3 * RB * Nz);   // the z roots, staged once
  for (int t = threadIdx.x; t < Nz; t += blockDim.x) tw[t] = gld(twz + t);
  const int j0 = blockIdx.x * RB, i = blockIdx.y;   // consecutive blocks: consecutive y of one kx
  const int nr = min(RB, g.my - j0);                 // rows of this block
  const long fstride = (long)g.my * g.kxn * Nz;
  const int nload = g.nkzb * nr * SNF_CW;
  const float rnr = 1.0f / (float)nr, rnz = 1.0f / (float)Nz;
  for (int grp = 0; grp < 2; ++grp) {
    for (int t = threadIdx.x; t < 3 * nr * Nz; t += blockDim.x) a[t] = wf_mk((R)0, (R)0);
    __syncthreads();
    for (int t = threadIdx.x; t < nload; t += blockDim.x) {
      const int c = t & (SNF_CW - 1), rc = t / SNF_CW, kc = xz_div(rc, rnr), r = rc - kc * nr;
      const int k = kc * SNF_CW + c;
      if (k >= g.mz) continue;
      const int j = j0 + r;
      PlaneIn q = load_in(g, vor, phi, hg, hv, dvordy, j, i, k);
      cplx d[6];
      derived6(g, q, i, k, gldd(u00 + j), gldd(w00 + j), gldd(dudy + j), gldd(dwdy + j), vbulk, d);
      const int pos = kz_pos(k, g.nz, g.mz, Nz);
#pragma unroll
      for (int f = 0; f < 3; ++f) a[(3 * r + f) * Nz + pos] = to_h<H>(d[grp * 3 + f]);
    }
    __syncthreads();
    H* res = fft_lds<false>(a, b, 3 * nr, pz, tw, +1);
    for (int t = threadIdx.x; t < 3 * nr * Nz; t += blockDim.x) {
      const int rf = xz_div(t, rnz), z = t - rf * Nz, r = rf / 3, f = rf - 3 * r;
      hyb[(grp * 3 + f) * fstride + ((long)(j0 + r) * g.kxn + i) * Nz + z] = res[t];
    }
    __syncthreads();
  }
}
template <class H>
__global__ void __launch_bounds__(256)
k_xz_A(Geom g, const cplx* __restrict__ vor, const cplx* __restrict__ phi,
       const cplx* __restrict__ hg, const cplx* __restrict__ hv,
       const cplx* __restrict__ dvordy, const double* __restrict__ u00,
       const double* __restrict__ w00, const double* __restrict__ dudy,
       const double* __restrict__ dwdy, double vbulk, H* __restrict__ hyb, FFTPlan pz,
       const cplx* __restrict__ twz, const DevScalars* sc) {
  k_xz_A_body<H>(g, vor, phi, hg, hv, dvordy, u00, w00, dudy, dwdy, vbulk, hyb, pz, twz, sc);
}

// ---------------------------------------------------------------------------
// B: ZB z-rows per block: c2r in x, u x omega, r2c in x.  Reads and writes
// runs of ZB consecutive z values per (field, kx) of the [f][y][kx][z]
// hybrid layout.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long dbits(double v) {
  return (unsigned long long)__double_as_longlong(v);
}

// LDS of the Stockham x-row kernel: 3 ZB sequences of mgalx elements of es
// bytes, then the mgalx FP64 x roots
inline size_t xz_lds_B(int zb, int mgalx, size_t es) {
  return 3 * (size_t)zb * mgalx * es + (size_t)mgalx * sizeof(cplx);
}

// NT threads per block (256, 512 or 1024), ZB z-rows per block
template <int NT, class H>
__device__ __forceinline__ void k_xz_B_body(Geom g, const H* __restrict__ hin, H* __restrict__ hout, FFTPlan px,
       const cplx* __restrict__ twx, int rk, unsigned long long* umax_acc,
       double* __restrict__ phist, const DevScalars* sc, int nhist, int ZB) {
  typedef typename wf_real<H>::T R;
  if (sc->halted) return;
  extern __shared__ double4 xz_smem[];
  H* lds = reinterpret_cast<H*>(xz_smem);
  const int Mx = g.mgalx, Nz = g.mgalz, kxn = g.kxn;
  // neighbouring z-blocks of one plane share 128-byte lines of the hybrid
  // field: keep them on one XCD
  const long lb = xcd_block((long)blockIdx.y * gridDim.x + blockIdx.x, (long)gridDim.x * gridDim.y);
  const int zb = (int)(lb % gridDim.x), j = (int)(lb / gridDim.x), z0 = ZB * zb;
  const long fstride = (long)g.my * kxn * Nz;
  const int ihist = ((sc->steps_local - 1) % nhist == 0) && rk == 1;
  const int nseq = 3 * ZB;    // sequence s = p*ZB + r (pair p, row r)
  // ZB is a power of two (2..16): row and group of an index by mask and shift
  const int zm = ZB - 1, zs = 31 - __clz(ZB);
  const float rmx = 1.0f / (float)Mx;   // t / Mx by xz_div
  cplx* tw = reinterpret_cast<cplx*>(lds + nseq * Mx);   // the x roots, staged once
  for (int t = threadIdx.x; t < Mx; t += blockDim.x) tw[t] = gld(twx + t);
  for (int t = threadIdx.x; t < nseq * Mx; t += blockDim.x) lds[t] = wf_mk((R)0, (R)0);
  __syncthreads();
  // Z_pair = G_a + i G_b with the Hermitian extension in kx
  for (int t = threadIdx.x; t < 3 * kxn * ZB; t += blockDim.x) {
    const int r = t & zm;
    const int pi_ = t >> zs;
    const int p = (pi_ >= kxn) + (pi_ >= 2 * kxn), i = pi_ - p * kxn;
    const long off = ((long)j * kxn + i) * Nz + z0 + r;
    const H Ga = hin[(2 * p) * fstride + off];
    const H Gb = hin[(2 * p + 1) * fstride + off];
    H* s = lds + (p * ZB + r) * Mx;
    if (i == 0) {
      s[0] = wf_mk(Ga.x, Gb.x);
    } else {
      s[i] = wf_mk(Ga.x - Gb.y, Ga.y + Gb.x);
      s[Mx - i] = wf_mk(Ga.x + Gb.y, Gb.x - Ga.y);
    }
  }
  __syncthreads();
  fft_lds_ip<XZ_MAXV>(lds, nseq, px, tw, +1);
  // u x omega: each (row pair, x) handled by one thread; results go to
  // sequences 0..ZB-1 (H1 + i H2 per row) and ZB..3ZB/2-1 (H3 of two rows)
  // products and maxima in the element type, the 'p' sums in FP64
  const R c2 = (R)(2.0 * g.chi);
  const R sy = (R)(g.s * ((double)j * g.dy - 0.5 * g.Ly));
  R m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  double p0 = 0.0, p1 = 0.0, p2 = 0.0;
  const int npair = ZB / 2;
  const int nwork = npair * Mx;
  constexpr int MAXW = XZ_B_MAXW;   // host-checked
  R h1[MAXW][2], h2[MAXW][2], h3[MAXW][2];
#pragma unroll
  for (int q = 0; q < MAXW; ++q) {
    const int t = threadIdx.x + q * blockDim.x;
    if (t < nwork) {
      const int rp = xz_div(t, rmx), x = t - rp * Mx;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = 2 * rp + h;
        const H z0v = lds[(0 * ZB + r) * Mx + x];
        const H z1v = lds[(1 * ZB + r) * Mx + x];
        const H z2v = lds[(2 * ZB + r) * Mx + x];
        const R u = z0v.x, v = z0v.y, w = z1v.x, o1 = z1v.y, o2 = z2v.x, o3 = z2v.y;
        if (rk == 1) {
          m0 = fmax(m0, fabs(u));
          m1 = fmax(m1, fabs(u + sy));
          m2 = fmax(m2, fabs(v));
          m3 = fmax(m3, fabs(w));
        }
        if (ihist) {
          p0 += (double)(v * (u * u));
          p1 += (double)(v * (w * w));
          p2 += (double)(v * v * v);
        }
        h1[q][h] = v * (o3 + c2) - w * o2;
        h2[q][h] = w * o1 - u * (o3 + c2);
        h3[q][h] = u * o2 - v * o1;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXW; ++q) {
    const int t = threadIdx.x + q * blockDim.x;
    if (t < nwork) {
      const int rp = xz_div(t, rmx), x = t - rp * Mx;
      lds[(2 * rp) * Mx + x] = wf_mk(h1[q][0], h2[q][0]);
      lds[(2 * rp + 1) * Mx + x] = wf_mk(h1[q][1], h2[q][1]);
      lds[(ZB + rp) * Mx + x] = wf_mk(h3[q][0], h3[q][1]);
    }
  }
  // block reductions: wave shuffles, then one value per wave (fixed order)
  __shared__ double red[4][NT / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  if (rk == 1) {
    for (int o = 32; o > 0; o >>= 1) {
      m0 = fmax(m0, __shfl_xor(m0, o, 64));
      m1 = fmax(m1, __shfl_xor(m1, o, 64));
      m2 = fmax(m2, __shfl_xor(m2, o, 64));
      m3 = fmax(m3, __shfl_xor(m3, o, 64));
    }
    if (lane == 0) {
      red[0][wv] = (double)m0;
      red[1][wv] = (double)m1;
      red[2][wv] = (double)m2;
      red[3][wv] = (double)m3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      double m = 0.0;
      for (int q = 0; q < nwv; ++q) m = fmax(m, red[threadIdx.x][q]);
      atomicMax(&umax_acc[threadIdx.x], dbits(m));
    }
    __syncthreads();
  }
  if (ihist) {
    for (int o = 32; o > 0; o >>= 1) {
      p0 += __shfl_xor(p0, o, 64);
      p1 += __shfl_xor(p1, o, 64);
      p2 += __shfl_xor(p2, o, 64);
    }
    if (lane == 0) {
      red[0][wv] = p0;
      red[1][wv] = p1;
      red[2][wv] = p2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      double acc = 0.0;
      for (int q = 0; q < nwv; ++q) acc += red[threadIdx.x][q];
      phist[((long)j * (Nz / ZB) + zb) * 3 + threadIdx.x] = acc;
    }
  }
  __syncthreads();
  fft_lds_ip<XZ_MAXV>(lds, ZB + npair, px, tw, -1);
  // unpack the pairs, z fastest among consecutive threads
  for (int t = threadIdx.x; t < kxn * ZB; t += blockDim.x) {
    const int r = t & zm, k = t >> zs;
    const int km = (k == 0) ? 0 : Mx - k;
    const H A = lds[r * Mx + k], Am = lds[r * Mx + km];
    const H* Bs = lds + (ZB + (r >> 1)) * Mx;
    const H B = Bs[k], Bm = Bs[km];
    const R hf = (R)0.5;
    const H H1 = wf_mk(hf * (A.x + Am.x), hf * (A.y - Am.y));
    const H H2 = wf_mk(hf * (A.y + Am.y), -hf * (A.x - Am.x));
    const H H3 = (r % 2 == 0) ? wf_mk(hf * (B.x + Bm.x), hf * (B.y - Bm.y))
                              : wf_mk(hf * (B.y + Bm.y), -hf * (B.x - Bm.x));
    const long off = ((long)j * kxn + k) * Nz + z0 + r;
    hout[off] = H1;
    hout[fstride + off] = H2;
    hout[2 * fstride + off] = H3;
  }
}
template <int NT, class H>
__global__ void __launch_bounds__(NT)
k_xz_B(Geom g, const H* __restrict__ hin, H* __restrict__ hout, FFTPlan px,
       const cplx* __restrict__ twx, int rk, unsigned long long* umax_acc,
       double* __restrict__ phist, const DevScalars* sc, int nhist, int ZB) {
  k_xz_B_body<NT, H>(g, hin, hout, px, twx, rk, umax_acc, phist, sc, nhist, ZB);
}

// ---------------------------------------------------------------------------
// C: hybrid -> spectral (forward z DFT), output combinations (iadd_mode=5,
// cross.f90:1533-1544) and the 00-mode terms rf0u, rf0w (cross.f90:1489-1490)
// ---------------------------------------------------------------------------
template <class H>
__device__ __forceinline__ void k_xz_C_body(Geom g, const H* __restrict__ hout, const cplx* __restrict__ vel,
       cplx* __restrict__ dvordy, cplx* __restrict__ hg, cplx* __restrict__ hv,
       double* __restrict__ rf0u,
       double* __restrict__ rf0w, FFTPlan pz, const cplx* __restrict__ twz,
       const DevScalars* sc) {
  if (sc->halted) return;
  extern __shared__ double4 xz_smem[];
  const int Nz = g.mgalz, RB = xz_rows_per_block(Nz);
  H* a = reinterpret_cast<H*>(xz_smem);
  H* b = a + 3 * RB * Nz;
  cplx* tw = reinterpret_cast<cplx*>(b + 3 * RB * Nz);   // the z roots, staged once
  for (int t = threadIdx.x; t < Nz; t += blockDim.x) tw[t] = gld(twz + t);
  const int j0 = blockIdx.x * RB, i = blockIdx.y;   // consecutive blocks: consecutive y of one kx
  const int nr = min(RB, g.my - j0);
  const long fstride = (long)g.my * g.kxn * Nz;
  const float rnr = 1.0f / (float)nr, rnz = 1.0f / (float)Nz;
  for (int t = threadIdx.x; t < 3 * nr * Nz; t += blockDim.x) {
    const int rf = xz_div(t, rnz), z = t - rf * Nz, r = rf / 3, f = rf - 3 * r;
    a[t] = hout[f * fstride + ((long)(j0 + r) * g.kxn + i) * Nz + z];
  }
  __syncthreads();
  H* res = fft_lds(a, b, 3 * nr, pz, tw, -1);
  const double norm = 1.0 / ((double)g.mgalx * (double)g.mgalz);
  const double kx = g.alp * (double)i;
  const double kx2 = kx * kx;
  const int nst = g.nkzb * nr * SNF_CW;
  for (int t = threadIdx.x; t < nst; t += blockDim.x) {
    const int c = t & (SNF_CW - 1), rc = t / SNF_CW, kc = xz_div(rc, rnr), r = rc - kc * nr;
    const int k = kc * SNF_CW + c;
    if (k >= g.mz) continue;
    const int j = j0 + r;
    const H* rr = res + 3 * r * Nz;
    const int pos = kz_pos(k, g.nz, g.mz, Nz);
    const cplx h1 = crmul(norm, to_d(rr[pos]));
    const cplx h2 = crmul(norm, to_d(rr[Nz + pos]));
    const cplx h3 = crmul(norm, to_d(rr[2 * Nz + pos]));
    const double kz = kz_of(k, g.nz, g.mz, g.gam);
    const long idx = sidx(g, j, i, k);
    const cplx v = ld_io(vel, idx, g.io32 & 1);
    st_io(dvordy, idx, cadd(mul_i(kx, h1), mul_i(kz, h3)), g.io32 & 2);
    const cplx d = mkc(h1.x - g.s * v.x, h1.y - g.s * v.y);
    st_io(hg, idx, cadd(mul_i(-kx, h3), mul_i(kz, d)), g.io32 & 2);
    const double kk = kx2 + kz * kz;
    st_io(hv, idx, mkc(-h2.x * kk, -h2.y * kk), g.io32 & 2);
    if (i == 0 && k == 0) {
      rf0u[j] = h1.x;
      rf0w[j] = h3.x;
    }
  }
}
template <class H>
__global__ void __launch_bounds__(256)
k_xz_C(Geom g, const H* __restrict__ hout, const cplx* __restrict__ vel,
       cplx* __restrict__ dvordy, cplx* __restrict__ hg, cplx* __restrict__ hv,
       double* __restrict__ rf0u,
       double* __restrict__ rf0w, FFTPlan pz, const cplx* __restrict__ twz,
       const DevScalars* sc) {
  k_xz_C_body<H>(g, hout, vel, dvordy, hg, hv, rf0u, rf0w, pz, twz, sc);
}

// ---------------------------------------------------------------------------
// hist 'f' (cross.f90:1798-1966): per-plane Fourier-space statistics.
// Output per plane j (STAT_W doubles):
//  0 up 1 vp 2 wp 3 uvr 4 uwr 5 vwr 6 o1p 7 o2p 8 o3p 9 o12p 10 eps
//  11 um 12 vm 13 wm 14 o1m 15 o2m 16 o3m
// ---------------------------------------------------------------------------
// Spectra of the 'f' statistics (cross.f90:1970-2016): for every (kx, kz)
// line, sums over the planes of |u|^2, |v|^2, |w|^2, Re(u v*), |o1|^2,
// |o2|^2, |o3|^2, Re(o1 o2*) (kx > 0 counted twice), added to sp and, times
// the step, to sp2 ([8][P]).  A block takes 64 consecutive lines; its
// SPEC_SEG waves walk fixed segments of the planes in order, and the segment
// sums are added in segment order, so the sums do not depend on the launch.
#define SPEC_SEG 8
__device__ __forceinline__ void k_hist_spec_body(Geom g, const cplx* __restrict__ vor, const cplx* __restrict__ phi,
            const cplx* __restrict__ hg, const cplx* __restrict__ hv,
            const cplx* __restrict__ dvordy, const double* __restrict__ u00,
            const double* __restrict__ w00, const double* __restrict__ dudy,
            const double* __restrict__ dwdy, double vbulk, double* __restrict__ sp,
            double* __restrict__ sp2, const DevScalars* sc, int nhist) {
  if (sc->halted) return;
  if ((sc->steps_local - 1) % nhist != 0) return;
  __shared__ double red[SPEC_SEG][8][64];
  const int lane = threadIdx.x & 63, seg = threadIdx.x >> 6;
  const long L = (long)blockIdx.x * 64 + lane;
  const bool on = L < g.P;
  const int i = on ? (int)(L / g.mz) : 0, k = on ? (int)(L - (long)i * g.mz) : 0;
  const double wt = (i == 0) ? 1.0 : 2.0;
  double acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = 0.0;
  const int j0 = (int)((long)g.my * seg / SPEC_SEG), j1 = on ? (int)((long)g.my * (seg + 1) / SPEC_SEG) : j0;
  for (int j = j0; j < j1; ++j) {
    PlaneIn q = load_in(g, vor, phi, hg, hv, dvordy, j, i, k);
    cplx d[6];
    derived6(g, q, i, k, u00[j], w00[j], dudy[j], dwdy[j], vbulk, d);
    acc[0] += wt * cabs2(d[0]);
    acc[1] += wt * cabs2(d[1]);
    acc[2] += wt * cabs2(d[2]);
    acc[3] += wt * cmulc(d[0], d[1]).x;
    acc[4] += wt * cabs2(d[3]);
    acc[5] += wt * cabs2(d[4]);
    acc[6] += wt * cabs2(d[5]);
    acc[7] += wt * cmulc(d[3], d[4]).x;
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) red[seg][q][lane] = acc[q];
  __syncthreads();
  if (seg == 0 && on) {
    const double dt = sc->Deltat;   // the step of the sample, as in the reference's 'f' pass
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      double t = red[0][q][lane];
#pragma unroll
      for (int s2 = 1; s2 < SPEC_SEG; ++s2) t += red[s2][q][lane];
      sp[q * g.P + L] += t;
      sp2[q * g.P + L] += t * dt;
    }
  }
}
__global__ void __launch_bounds__(64 * SPEC_SEG)
k_hist_spec(Geom g, const cplx* __restrict__ vor, const cplx* __restrict__ phi,
            const cplx* __restrict__ hg, const cplx* __restrict__ hv,
            const cplx* __restrict__ dvordy, const double* __restrict__ u00,
            const double* __restrict__ w00, const double* __restrict__ dudy,
            const double* __restrict__ dwdy, double vbulk, double* __restrict__ sp,
            double* __restrict__ sp2, const DevScalars* sc, int nhist) {
  k_hist_spec_body(g, vor, phi, hg, hv, dvordy, u00, w00, dudy, dwdy, vbulk, sp, sp2, sc, nhist);
}

__device__ __forceinline__ void k_hist_f_body(Geom g, const cplx* __restrict__ vor, const cplx* __restrict__ phi,
         const cplx* __restrict__ hg, const cplx* __restrict__ hv,
         const cplx* __restrict__ dvordy, const double* __restrict__ u00,
         const double* __restrict__ w00, const double* __restrict__ dudy,
         const double* __restrict__ dwdy, double vbulk, double* __restrict__ out,
         const DevScalars* sc, int nhist) {
  if (sc->halted) return;
  if ((sc->steps_local - 1) % nhist != 0) return;
  const int j = blockIdx.x;
  double acc[11];
#pragma unroll
  for (int q = 0; q < 11; ++q) acc[q] = 0.0;
  for (long L = threadIdx.x; L < g.P; L += blockDim.x) {
    const int i = (int)(L / g.mz), k = (int)(L - (long)i * g.mz);
    PlaneIn q = load_in(g, vor, phi, hg, hv, dvordy, j, i, k);
    cplx d[6];
    derived6(g, q, i, k, u00[j], w00[j], dudy[j], dwdy[j], vbulk, d);
    const double wt = (i == 0) ? 1.0 : 2.0;
    const cplx u = d[0], v = d[1], w = d[2], o1 = d[3], o2 = d[4];
    cplx o3 = d[5];
    acc[0] += wt * cabs2(u);
    acc[1] += wt * cabs2(v);
    acc[2] += wt * cabs2(w);
    acc[3] += wt * cmulc(u, v).x;
    acc[4] += wt * cmulc(u, w).x;
    acc[5] += wt * cmulc(v, w).x;
    acc[6] += wt * cabs2(o1);
    acc[7] += wt * cabs2(o2);
    acc[8] += wt * cabs2(o3);
    acc[9] += wt * cmulc(o1, o2).x;
    // dissipation, with the mean shear removed from o3(0,0)
    const double kx = g.alp * (double)i;
    const double kz = kz_of(k, g.nz, g.mz, g.gam);
    if (i == 0 && k == 0) o3.x -= g.s;
    const double uu = cabs2(u) + cabs2(v) + cabs2(w);
    const cplx c1 = cadd(o1, mul_i(kz, v));
    if (i == 0) {
      acc[10] += kz * kz * uu + cabs2(q.dv) + cabs2(o3) + cabs2(c1);
    } else {
      const cplx c3 = csub(o3, mul_i(kx, v));
      acc[10] += 2.0 * ((kx * kx + kz * kz) * uu + cabs2(q.dv) + cabs2(c1) + cabs2(c3));
    }
    if (L == 0) {
      out[(long)j * STAT_W + 11] = u.x;
      out[(long)j * STAT_W + 12] = v.x;
      out[(long)j * STAT_W + 13] = w.x;
      out[(long)j * STAT_W + 14] = o1.x;
      out[(long)j * STAT_W + 15] = o2.x;
      out[(long)j * STAT_W + 16] = d[5].x;
    }
  }
  __shared__ double red[11][256];
#pragma unroll
  for (int q = 0; q < 11; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
#pragma unroll
      for (int q = 0; q < 11; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x < 11) out[(long)j * STAT_W + threadIdx.x] = red[threadIdx.x][0];
}
__global__ void __launch_bounds__(256)
k_hist_f(Geom g, const cplx* __restrict__ vor, const cplx* __restrict__ phi,
         const cplx* __restrict__ hg, const cplx* __restrict__ hv,
         const cplx* __restrict__ dvordy, const double* __restrict__ u00,
         const double* __restrict__ w00, const double* __restrict__ dudy,
         const double* __restrict__ dwdy, double vbulk, double* __restrict__ out,
         const DevScalars* sc, int nhist) {
  k_hist_f_body(g, vor, phi, hg, hv, dvordy, u00, w00, dudy, dwdy, vbulk, out, sc, nhist);
}

// ---------------------------------------------------------------------------
// standalone fourxz for the operator ABI: planes [p][kx][kz] <-> [p][x][z]
// ---------------------------------------------------------------------------
__global__ void k_fourxz_fwd_z(Geom g, const cplx* in, cplx* hyb, FFTPlan pz, const cplx* twz) {
  // spectral -> hybrid [p][kx][z] (inverse z DFT)
  extern __shared__ cplx lds[];
  const int Nz = g.mgalz;
  cplx* a = lds;
  cplx* b = lds + Nz;
  const int i = blockIdx.x, p = blockIdx.y;
  for (int t = threadIdx.x; t < Nz; t += blockDim.x) a[t] = mkc(0.0, 0.0);
  __syncthreads();
  for (int k = threadIdx.x; k < g.mz; k += blockDim.x)
    a[kz_pos(k, g.nz, g.mz, Nz)] = in[((long)p * g.kxn + i) * g.mz + k];
  __syncthreads();
  cplx* r = fft_lds(a, b, 1, pz, twz, +1);
  for (int z = threadIdx.x; z < Nz; z += blockDim.x) hyb[((long)p * g.kxn + i) * Nz + z] = r[z];
}

__global__ void k_fourxz_x_c2r(Geom g, const cplx* hyb, double* out, FFTPlan px, const cplx* twx) {
  // one z row per block: out[p][x][z]
  extern __shared__ cplx lds[];
  const int Mx = g.mgalx, Nz = g.mgalz;
  cplx* a = lds;
  cplx* b = lds + Mx;
  const int z = blockIdx.x, p = blockIdx.y;
  for (int t = threadIdx.x; t < Mx; t += blockDim.x) a[t] = mkc(0.0, 0.0);
  __syncthreads();
  for (int i = threadIdx.x; i < g.kxn; i += blockDim.x) {
    const cplx G = hyb[((long)p * g.kxn + i) * Nz + z];
    if (i == 0) {
      a[0] = mkc(G.x, 0.0);
    } else {
      a[i] = G;
      a[Mx - i] = cconj(G);
    }
  }
  __syncthreads();
  cplx* r = fft_lds(a, b, 1, px, twx, +1);
  for (int x = threadIdx.x; x < Mx; x += blockDim.x) out[((long)p * Mx + x) * Nz + z] = r[x].x;
}

__global__ void k_fourxz_x_r2c(Geom g, const double* in, cplx* hyb, FFTPlan px, const cplx* twx) {
  extern __shared__ cplx lds[];
  const int Mx = g.mgalx, Nz = g.mgalz;
  cplx* a = lds;
  cplx* b = lds + Mx;
  const int z = blockIdx.x, p = blockIdx.y;
  for (int x = threadIdx.x; x < Mx; x += blockDim.x) a[x] = mkc(in[((long)p * Mx + x) * Nz + z], 0.0);
  __syncthreads();
  cplx* r = fft_lds(a, b, 1, px, twx, -1);
  for (int i = threadIdx.x; i < g.kxn; i += blockDim.x) hyb[((long)p * g.kxn + i) * Nz + z] = r[i];
}

__global__ void k_fourxz_inv_z(Geom g, const cplx* hyb, cplx* out, FFTPlan pz, const cplx* twz) {
  extern __shared__ cplx lds[];
  const int Nz = g.mgalz;
  cplx* a = lds;
  cplx* b = lds + Nz;
  const int i = blockIdx.x, p = blockIdx.y;
  for (int z = threadIdx.x; z < Nz; z += blockDim.x) a[z] = hyb[((long)p * g.kxn + i) * Nz + z];
  __syncthreads();
  cplx* r = fft_lds(a, b, 1, pz, twz, -1);
  const double norm = 1.0 / ((double)g.mgalx * (double)g.mgalz);
  for (int k = threadIdx.x; k < g.mz; k += blockDim.x)
    out[((long)p * g.kxn + i) * g.mz + k] = crmul(norm, r[kz_pos(k, g.nz, g.mz, Nz)]);
}
