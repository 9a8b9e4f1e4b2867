// Box-table kernels of the batched ensemble: one launch of a kernel advances
// every box of a size class (boxes with the same grid, precision and viscous
// mode, which take the same kernels with the same launch shapes).  The box
// is blockIdx.z; its arguments (fields, scalars, Reynolds number, twist) are
// row blockIdx.z of a device table that the host fills per substep.  Each
// kernel runs the same body as its single-box form, so a batched box is
// bitwise the box advanced alone.
//
// The table path covers the sizes of the ensemble workload: the Stockham x-z
// kernels (planes other than 64*R points) and the spectral y stage (line
// lengths up to 144 or with a radix plan), without y-dealiasing.
#include "snf_internal.h"

// the real copies of U, W that the x-z stage reads
__device__ __forceinline__ void k_cplx_to_real_body(const cplx* a, const cplx* b, double* ra,
                                                    double* rb, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) {
    ra[j] = a[j].x;
    rb[j] = b[j].x;
  }
}
__global__ void k_cplx_to_real(const cplx* a, const cplx* b, double* ra, double* rb, int n) {
  k_cplx_to_real_body(a, b, ra, rb, n);
}

// ---- one box's arguments of each batched kernel
struct BxStep {
  DevScalars* sc;
  double wrap_lx;
};
struct BxC2R {
  const cplx *a, *b;
  double *ra, *rb;
  int n;
};
struct BxXzA {
  Geom g;
  const cplx *vor, *phi, *v, *dv, *dome;
  const double *u00r, *w00r, *dudy, *dwdy;
  double vbulk;
  void* hyb;
  FFTPlan pz;
  const cplx* twz;
  const DevScalars* sc;
};
struct BxHist {   // k_hist_f (out) and k_hist_spec (out = sp, out2 = sp2)
  Geom g;
  const cplx *vor, *phi, *v, *dv, *dome;
  const double *u00r, *w00r, *dudy, *dwdy;
  double vbulk;
  double *out, *out2;
  const DevScalars* sc;
  int nhist;
};
struct BxXzB {
  Geom g;
  const void* hin;
  void* hout;
  FFTPlan px;
  const cplx* twx;
  unsigned long long* umax_acc;
  double* phist;
  const DevScalars* sc;
  int nhist, zb;
};
struct BxXzC {
  Geom g;
  const void* hout;
  const cplx* vel;
  cplx *dvordy, *hg, *hv;
  double *rf0u, *rf0w;
  FFTPlan pz;
  const cplx* twz;
  const DevScalars* sc;
};
struct BxPhist {
  const double* phist;
  int nzb;
  double* pplane;
  const DevScalars* sc;
  int nhist;
};
struct BxYsym {
  int N;
  YCoef c;
  const cplx *shp, *rt;
  cplx* out;
  const DevScalars* sc;
};

// ---- table kernels: row blockIdx.z of the table is the box
__global__ void k_step_begin_tab(const BxStep* __restrict__ t) {
  const BxStep& a = t[blockIdx.z];
  k_step_begin_body(a.sc, a.wrap_lx);
}
__global__ void k_cplx_to_real_tab(const BxC2R* __restrict__ t) {
  const BxC2R& a = t[blockIdx.z];
  k_cplx_to_real_body(a.a, a.b, a.ra, a.rb, a.n);
}
template <class H>
__global__ void __launch_bounds__(256) k_xz_A_tab(const BxXzA* __restrict__ t) {
  const BxXzA& a = t[blockIdx.z];
  k_xz_A_body<H>(a.g, a.vor, a.phi, a.v, a.dv, a.dome, a.u00r, a.w00r, a.dudy, a.dwdy, a.vbulk,
                 reinterpret_cast<H*>(a.hyb), a.pz, a.twz, a.sc);
}
__global__ void __launch_bounds__(256) k_hist_f_tab(const BxHist* __restrict__ t) {
  const BxHist& a = t[blockIdx.z];
  k_hist_f_body(a.g, a.vor, a.phi, a.v, a.dv, a.dome, a.u00r, a.w00r, a.dudy, a.dwdy, a.vbulk,
                a.out, a.sc, a.nhist);
}
__global__ void __launch_bounds__(64 * SPEC_SEG) k_hist_spec_tab(const BxHist* __restrict__ t) {
  const BxHist& a = t[blockIdx.z];
  k_hist_spec_body(a.g, a.vor, a.phi, a.v, a.dv, a.dome, a.u00r, a.w00r, a.dudy, a.dwdy, a.vbulk,
                   a.out, a.out2, a.sc, a.nhist);
}
template <int NT, class H>
__global__ void __launch_bounds__(NT) k_xz_B_tab(const BxXzB* __restrict__ t, int rk) {
  const BxXzB& a = t[blockIdx.z];
  k_xz_B_body<NT, H>(a.g, reinterpret_cast<const H*>(a.hin), reinterpret_cast<H*>(a.hout), a.px,
                     a.twx, rk, a.umax_acc, a.phist, a.sc, a.nhist, a.zb);
}
template <int RZ, class H>
__global__ void __launch_bounds__(256) k_xzs_A_tab(const BxXzA* __restrict__ t) {
  const BxXzA& a = t[blockIdx.z];
  k_xzs_A_body<RZ, H>(a.g, a.vor, a.phi, a.v, a.dv, a.dome, a.u00r, a.w00r, a.dudy, a.dwdy, a.vbulk,
                      reinterpret_cast<H*>(a.hyb), a.twz, a.sc, (int)xcd_block(blockIdx.x, gridDim.x));
}
template <int RZ, class H>
__global__ void __launch_bounds__(256) k_xzs_C_tab(const BxXzC* __restrict__ t) {
  const BxXzC& a = t[blockIdx.z];
  k_xzs_C_body<RZ, H>(a.g, reinterpret_cast<const H*>(a.hout), a.vel, a.dvordy, a.hg, a.hv, a.rf0u,
                      a.rf0w, a.twz, a.sc, (int)xcd_block(blockIdx.x, gridDim.x));
}
template <int RX, class H>
__global__ void __launch_bounds__(256) k_xzs_B_tab(const BxXzB* __restrict__ t, int rk) {
  const BxXzB& a = t[blockIdx.z];
  const long lb = xcd_block((long)blockIdx.y * gridDim.x + blockIdx.x, (long)gridDim.x * gridDim.y);
  k_xzs_B_body<RX, H>(a.g, reinterpret_cast<const H*>(a.hin), reinterpret_cast<H*>(a.hout), a.twx,
                      rk, a.umax_acc, a.phist, a.sc, a.nhist, (int)(lb % gridDim.x),
                      (int)(lb / gridDim.x));
}
template <class H>
__global__ void __launch_bounds__(256) k_xz_C_tab(const BxXzC* __restrict__ t) {
  const BxXzC& a = t[blockIdx.z];
  k_xz_C_body<H>(a.g, reinterpret_cast<const H*>(a.hout), a.vel, a.dvordy, a.hg, a.hv, a.rf0u,
                 a.rf0w, a.pz, a.twz, a.sc);
}
__global__ void __launch_bounds__(64) k_phist_plane_tab(const BxPhist* __restrict__ t) {
  const BxPhist& a = t[blockIdx.z];
  k_phist_plane_body(a.phist, a.nzb, a.pplane, a.sc, a.nhist);
}
__global__ void __launch_bounds__(256) k_scalar_tab(const ScalarArgs* __restrict__ t, int rk) {
  k_scalar_body(t[blockIdx.z], rk);
}
__global__ void __launch_bounds__(256) k_ysym_tab(const BxYsym* __restrict__ t) {
  const BxYsym& a = t[blockIdx.z];
  k_ysym_body(a.N, a.c, a.shp, a.rt, a.out, a.sc);
}

