// Spectral y plans: the direct DFT up to 192 points, and on the matrix cores up
// to 128 points.
#include "snf_yplan_impl.h"

SNF_YPLAN(direct3, YDirect<3>)
SNF_YPLAN(mfma1, YMfma<1>)
SNF_YPLAN(mfma2, YMfma<2>)
