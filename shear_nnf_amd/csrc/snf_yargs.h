// Arguments of the y-line kernels and the per-substep RK3 factors, shared by
// the banded y kernels and the spectral y-plan translation units.
#pragma once
#include "snf_internal.h"
#include "snf_lines.h"
#include "snf_rk.h"

// ---------------------------------------------------------------------------
// arguments of the y-line kernels
// ---------------------------------------------------------------------------
struct YArgs {
  Geom g;
  YCoef c;
  BaseFactor b1, b2;
  TwistTail id1, id2;                       // untwisted tails (00 lines)
  const TwistTail* tw_old;                  // [2][kxn] D1, D2 at the current twist
  const TwistTail* tw_new;                  // [2][kxn] after the shift update
  const cplx *shp_old, *shm_old, *shp_new, *shm_new;
  cplx *vor, *phi, *vorwk, *phiwk;          // state and RK accumulators
  cplx *v, *dome, *dv;                      // v, d(omega_y)/dy, dv/dy for the x-z stage
  cplx *rdome, *rhg, *rhv;                  // x-z stage outputs (cross.f90:1533-1544)
  cplx *s1, *s2, *s3;                       // forward-sweep scratch, stride PS
  HScratch hs;                              // Helmholtz factor scratch, stride PS
  cplx *u00c, *w00c, *u00wk, *w00wk;        // 00 lines (real data, imag 0), stride 1
  const double *rf0u, *rf0w;
  double *dudy, *dwdy;
  DevScalars* sc;
  long PS;
  const cplx* twy;                          // spectral y path: exp(-2 pi i t/my), t < my
  const cplx* tpy;                          // ... pass-major twiddles of the transform plan
  const cplx *ysym_old, *ysym_new;          // [kxn][my][2] multipliers of d/dy, d2/dy2
};

// dt and the moving-frame factors f1, f2 of substep rk (cross.f90:652-781)
__device__ __forceinline__ void substep_factors(const DevScalars* sc, double s, int rk,
                                                double& dt, double& f1, double& f2) {
  dt = sc->Deltat;
  const int ir = rk - 1;
  f1 = (rk == 1) ? -(c_gama[0] * dt * s) : -((c_gama[1] + c_zeta[0]) * dt * s);
  const double r1 = c_ca[ir] * dt;
  f2 = (rk == 1) ? -(s * r1) : -(s * (c_ca[ir] - c_ca[ir - 1]) * dt);
}
