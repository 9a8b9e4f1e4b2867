// Spectral y plans: the direct DFT on the matrix cores up to 192 points, and the
// direct-DFT compact filter of line lengths without a spectral plan.
#include "snf_yplan_impl.h"

SNF_YPLAN(mfma3, YMfma<3>)
SNF_YFILTER_DIRECT(fdirect5, YDirect<5>)
SNF_YFILTER_DIRECT(fdirect9, YDirect<9>)
SNF_YFILTER_DIRECT(fdirect17, YDirect<17>)
