// Host side of the spectral y stage, one table of entry points per line
// plan.  Each plan's kernels (snf_yfk.h, instantiated in snf_yplan_impl.h)
// are compiled in a translation unit of their own (snf_yplans_*.hip), so the
// plans build in parallel; the library finds a plan's table by line length.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

struct YArgs;

// what a launch takes from the handle
struct YLaunch {
  hipStream_t stream;
  size_t lds;          // dynamic LDS of the plan's kernels (bytes per block)
  int explicit_visc;   // 1: explicit viscous terms, 0: implicit
  int io32;            // storage of the inter-stage fields (Geom::io32)
  int y32;             // mixed mode: explicit substeps with FP32 line transforms
};

struct YPlanOps {
  // dynamic LDS of the plan's kernels at line length my (bytes per block)
  size_t (*lds_bytes)(int my);
  // one substep of the y stage (explicit or implicit viscous terms)
  void (*substep)(const YLaunch& c, unsigned nb, const YArgs& a, int rk);
  // start of a cross call: v, dv, dome, dU/dy, dW/dy at the current twist
  void (*first)(const YLaunch& c, unsigned nb, const YArgs& a);
  // compact filter of vor, phi, U, W at the current twist
  void (*filter)(const YLaunch& c, unsigned nb, const YArgs& a);
  // one substep of n boxes from a device table of arguments (n = 0: the
  // LDS limits only, before a graph capture)
  void (*table)(const YLaunch& c, unsigned nb, const YArgs* t, int n, int rk, int dev);
  // raise the LDS limit of every kernel of the plan to lds on device dev
  void (*lds_attrs)(size_t lds, int dev);
};

// the spectral plan of line length my (the direct DFT on the matrix cores
// when mfma), or nullptr where my has none
const YPlanOps* yplan_ops(int my, bool mfma);
// the direct-DFT filter of line lengths without a spectral plan (the
// banded substeps): only lds_bytes, filter and lds_attrs are set
const YPlanOps* yplan_filter_direct_ops(int my);

// raise a kernel's LDS limit on device dev to at least bytes (never lowers it;
// defined in snf_api.hip)
void lds_attr(const void* k, size_t bytes, int dev);
