// The entry points of one spectral y plan PL (snf_yplan.h), instantiated
// by the plan translation units snf_yplans_*.hip through SNF_YPLAN.
#pragma once
#include "snf_yfk.h"
#include "snf_yfk32.h"
#include "snf_yplan.h"

template <class PL>
struct YPlanImpl {
  // the transfer tile and a column per wave, the plan's table, the two
  // multiplier rows (unless the plan reads them from the global table),
  // four phase tables: 81552 B at my = 385 (two blocks fit a CU's 160 KiB),
  // 75472 B at my = 513 with the rows read from the global table
  static size_t lds_bytes(int my) {
    return ((size_t)my * 2 * YF_LINES + 2 * (size_t)ys_lds<PL>(my) + (size_t)PL::tw_len(my) +
            4 * YPH_N((my + 63) / 64)) * sizeof(cplx);
  }
  template <int IO>
  static void substep_io(const YLaunch& c, unsigned nb, const YArgs& a, int rk) {
    if (c.explicit_visc == 1)
      k_yf_explicit<PL, IO><<<nb, 64 * YF_LINES, c.lds, c.stream>>>(a, rk);
    else
      k_yf_implicit<PL, IO><<<nb, 64 * YF_LINES, c.lds, c.stream>>>(a, rk);
  }
  // the mixed mode's explicit substep with FP32 transforms (radix plans)
  static bool use32(const YLaunch& c) {
    if constexpr (PL::f32) return c.y32 && c.io32 && c.explicit_visc == 1;
    return false;
  }
  static void substep(const YLaunch& c, unsigned nb, const YArgs& a, int rk) {
    if constexpr (PL::f32) {
      if (use32(c)) {
        k_yf_explicit32<PL><<<nb, 64 * YF_LINES, Y32<PL>::lds_bytes, c.stream>>>(a, rk);
        return;
      }
    }
    if (c.io32)
      substep_io<3>(c, nb, a, rk);
    else
      substep_io<0>(c, nb, a, rk);
  }
  static void first(const YLaunch& c, unsigned nb, const YArgs& a) {
    if (c.io32)
      k_yf_first<PL, 3><<<nb, 64 * YF_LINES, c.lds, c.stream>>>(a);
    else
      k_yf_first<PL, 0><<<nb, 64 * YF_LINES, c.lds, c.stream>>>(a);
  }
  static void filter(const YLaunch& c, unsigned nb, const YArgs& a) {
    k_yf_filter<PL><<<nb, 64 * YF_LINES, c.lds, c.stream>>>(a);
  }
  template <int IO>
  static void table_io(const YLaunch& c, unsigned nb, const YArgs* t, int n, int rk, int dev) {
    if (c.explicit_visc == 1) {
      lds_attr((const void*)k_yf_explicit_tab<PL, IO>, c.lds, dev);
      if (n == 0) return;
      k_yf_explicit_tab<PL, IO><<<dim3(nb, 1, n), 64 * YF_LINES, c.lds, c.stream>>>(t, rk);
    } else {
      lds_attr((const void*)k_yf_implicit_tab<PL, IO>, c.lds, dev);
      if (n == 0) return;
      k_yf_implicit_tab<PL, IO><<<dim3(nb, 1, n), 64 * YF_LINES, c.lds, c.stream>>>(t, rk);
    }
  }
  static void table(const YLaunch& c, unsigned nb, const YArgs* t, int n, int rk, int dev) {
    if constexpr (PL::f32) {
      if (use32(c)) {
        lds_attr((const void*)k_yf_explicit32_tab<PL>, Y32<PL>::lds_bytes, dev);
        if (n == 0) return;
        k_yf_explicit32_tab<PL><<<dim3(nb, 1, n), 64 * YF_LINES, Y32<PL>::lds_bytes, c.stream>>>(t, rk);
        return;
      }
    }
    if (c.io32)
      table_io<3>(c, nb, t, n, rk, dev);
    else
      table_io<0>(c, nb, t, n, rk, dev);
  }
  static void lds_attrs(size_t lds, int dev) {
    for (const void* k : {(const void*)k_yf_explicit<PL, 0>, (const void*)k_yf_implicit<PL, 0>,
                          (const void*)k_yf_first<PL, 0>, (const void*)k_yf_explicit<PL, 3>,
                          (const void*)k_yf_implicit<PL, 3>, (const void*)k_yf_first<PL, 3>,
                          (const void*)k_yf_filter<PL>})
      lds_attr(k, lds, dev);
    if constexpr (PL::f32) lds_attr((const void*)k_yf_explicit32<PL>, Y32<PL>::lds_bytes, dev);
  }
  static const YPlanOps* ops() {
    static const YPlanOps o = {lds_bytes, substep, first, filter, table, lds_attrs};
    return &o;
  }
};

// The direct-DFT compact filter of line lengths without a spectral plan, in
// an LDS plan of its own: the transfer tile, a column per wave, the roots.
template <class PL>
struct YFilterDirectImpl {
  static size_t lds_bytes(int my) { return (size_t)my * (2 * YF_LINES + 1) * sizeof(cplx); }
  static void filter(const YLaunch& c, unsigned nb, const YArgs& a) {
    k_yf_filter<PL><<<nb, 64 * YF_LINES, c.lds, c.stream>>>(a);
  }
  static void lds_attrs(size_t lds, int dev) { lds_attr((const void*)k_yf_filter<PL>, lds, dev); }
  static const YPlanOps* ops() {
    static const YPlanOps o = {lds_bytes, nullptr, nullptr, filter, nullptr, lds_attrs};
    return &o;
  }
};

// the entry point table getter of plan PL, named yplan_NAME
#define SNF_YPLAN(NAME, ...) \
  const YPlanOps* yplan_##NAME() { return YPlanImpl<__VA_ARGS__>::ops(); }
#define SNF_YFILTER_DIRECT(NAME, ...) \
  const YPlanOps* yplan_##NAME() { return YFilterDirectImpl<__VA_ARGS__>::ops(); }
