// Internal definitions shared by the HIP translation units of libshear_rk3.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double2 cplx;
typedef float2 cplxf;   // FP32 complex: x-z stage of the mixed precision mode

#define SNF_MAXN 1032      // largest my supported (line length)
#define SNF_MAXFFT 1024    // largest mgalx / mgalz
#define SNF_NSTAT 23
#define SNF_NSPEC 8        // spectra of the statistics (nspec)

__host__ __device__ __forceinline__ cplx mkc(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return mkc(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return mkc(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return mkc(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ cplx cmulc(cplx a, cplx b) {  // a * conj(b)
  return mkc(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ cplx crmul(double r, cplx a) { return mkc(r * a.x, r * a.y); }
__device__ __forceinline__ cplx cconj(cplx a) { return mkc(a.x, -a.y); }
__device__ __forceinline__ double cabs2(cplx a) { return a.x * a.x + a.y * a.y; }

// Loads and stores through the global address space.  Pointers that reach
// a device function through a struct lose their address space, and the
// generic (flat) accesses the compiler then emits are waited for with
// vmcnt(0) lgkmcnt(0), which serialises every prefetch behind them.
#define SNF_AS1 __attribute__((address_space(1)))
__device__ __forceinline__ cplx gld(const cplx* p) {
  const SNF_AS1 double* d = (const SNF_AS1 double*)p;
  return make_double2(d[0], d[1]);
}
__device__ __forceinline__ void gst(cplx* p, cplx v) {
  SNF_AS1 double* d = (SNF_AS1 double*)p;
  d[0] = v.x;
  d[1] = v.y;
}
__device__ __forceinline__ cplxf gld(const cplxf* p) {
  const SNF_AS1 float* d = (const SNF_AS1 float*)p;
  return make_float2(d[0], d[1]);
}
__device__ __forceinline__ void gst(cplxf* p, cplxf v) {
  SNF_AS1 float* d = (SNF_AS1 float*)p;
  d[0] = v.x;
  d[1] = v.y;
}
// conversions between the FP64 fields and an x-z element type H
template <class H> __device__ __forceinline__ H to_h(cplx a);
template <> __device__ __forceinline__ cplx to_h<cplx>(cplx a) { return a; }
template <> __device__ __forceinline__ cplxf to_h<cplxf>(cplx a) {
  return make_float2((float)a.x, (float)a.y);
}
__device__ __forceinline__ cplx to_d(cplx a) { return a; }
__device__ __forceinline__ cplx to_d(cplxf a) { return make_double2((double)a.x, (double)a.y); }
__device__ __forceinline__ double gldd(const double* p) { return *(const SNF_AS1 double*)p; }
__device__ __forceinline__ void gstd(double* p, double v) { *(SNF_AS1 double*)p = v; }

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  xcd_block renumbers a linear block index so that
// consecutive logical blocks run on the same XCD at about the same time and
// share the cache lines they both touch.
#define SNF_NXCD 8
__device__ __forceinline__ long xcd_block(long b, long nb) {
  const long q = nb / SNF_NXCD;
  if (b >= q * SNF_NXCD) return b;
  return (b % SNF_NXCD) * q + b / SNF_NXCD;
}

// kz of z-index k: 0..nz non-negative, nz+1..mz-1 negative
__host__ __device__ __forceinline__ double kz_of(int k, int nz, int mz, double gam) {
  return k <= nz ? gam * (double)k : -gam * (double)(mz - k);
}

// Grid description passed by value to every kernel.
struct Geom {
  int my, kxn, mz, nz, mgalx, mgalz;
  int nkzb;          // kz chunks of SNF_CW lines: ceil(mz / SNF_CW)
  long P;            // kxn*mz: complex coefficients per y-plane (= number of y-lines)
  double alp, gam, Ly, dy, s, chi, Re;
  // Storage of the fields between the two stages of a substep, bit 0: v,
  // dv/dy, d(omega_y)/dy (y stage -> x-z stage), bit 1: rdome, rhg, rhv
  // (x-z stage -> y stage); a set bit stores those three as FP32 complex
  // (the mixed precision mode with a spectral y stage).  Arithmetic on them
  // stays FP64.
  int io32;
};

// Device layout of every spectral field (vor, phi, the RK accumulators and
// the inter-stage fields): kz-chunk-major, [kx][kz / SNF_CW][y][kz % SNF_CW].
// The SNF_CW lines of a chunk are one contiguous run of my * SNF_CW elements,
// which the y stage moves whole; a row (kx, y) of the x-z stage is nkzb
// pieces of SNF_CW elements, which blocks of consecutive y read together.
// Line (i, k) starts at sline(g, i, k) and has stride SNF_CW in y.  A field
// holds kxn * nkzb * my * SNF_CW elements; those with k >= mz (the last chunk
// when SNF_CW does not divide mz) stay zero and are never read.
#define SNF_CW 4
__host__ __device__ __forceinline__ long sidx(const Geom& g, int j, int i, int k) {
  return (((long)i * g.nkzb + (k >> 2)) * g.my + j) * SNF_CW + (k & (SNF_CW - 1));
}
__host__ __device__ __forceinline__ long sline(const Geom& g, int i, int k) { return sidx(g, 0, i, k); }
__host__ __device__ __forceinline__ long sfield_len(const Geom& g) {
  return (long)g.kxn * g.nkzb * g.my * SNF_CW;
}
static_assert(SNF_CW == 4, "sidx takes k >> 2 as the chunk of k");

// loads and stores of the inter-stage fields in the storage the grid says
__device__ __forceinline__ cplx ld_io(const cplx* f, long idx, int io32) {
  return io32 ? to_d(gld(reinterpret_cast<const cplxf*>(f) + idx)) : gld(f + idx);
}
__device__ __forceinline__ void st_io(cplx* f, long idx, cplx v, int io32) {
  if (io32)
    gst(reinterpret_cast<cplxf*>(f) + idx, to_h<cplxf>(v));
  else
    gst(f + idx, v);
}

// Compact-scheme stencils (offsets -3..3) and untwisted base factors.
struct YCoef {
  double rh1[7], rh2[7], pr1[7], pr2[7];
};

// Untwisted (twist = 1) L D L^H factors of the first/second-derivative LHS;
// border rows kept as real numbers (their imaginary part is zero).
struct BaseFactor {
  const double *D, *Dm1, *La, *Lb, *Lc, *Ler, *Lfr, *Lgr;  // N entries each
};

// Per-kx re-factored tail of a twisted derivative factor (modChol_addshear):
// Le rows N-6..N-4, Lf rows N-5..N-3, Lg rows N-4..N-2, Dm1 rows N-3..N-1.
struct TwistTail {
  cplx Le[3], Lf[3], Lg[3];
  double Dm1[3];
};

// Device-resident run scalars (modules running / bcs).
struct DevScalars {
  double time, Deltat, dtr, xoffb, xofft, xwkb, xwkt;
  double etime, dumptime, dtimag;
  double umax[4], CFLv;
  int istop, idump, halted, steps_done;
  int steps_local;    // istep of the current advance call (1-based)
  int nrec;           // UNER records written (ring position)
  int nacum;
  double total_dt;
  int stop_reason;    // snf_stop_reason of the first istop of the run
};

// Derived per-plane statistics produced by the hist kernels:
// per plane j, STAT_W doubles (see snf_xz.hip k_hist_f for the layout).
#define STAT_W 24
