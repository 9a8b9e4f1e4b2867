// SMR RK3 constants (cross.f90:18-24), one copy per translation unit.
#pragma once
#include <hip/hip_runtime.h>

static __constant__ double c_gama[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static __constant__ double c_alpha[3] = {29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0};
static __constant__ double c_ibeta[3] = {160.0 / 37.0, 24.0 / 5.0, 6.0};
static __constant__ double c_zeta[3] = {-17.0 / 60.0, -5.0 / 12.0, 0.0};
static __constant__ double c_ca[3] = {8.0 / 15.0, 2.0 / 3.0, 1.0};
