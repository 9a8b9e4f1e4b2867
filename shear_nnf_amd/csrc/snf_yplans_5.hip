// Spectral y plans: line length 256, and the direct DFT up to 128 points.
#include "snf_yplan_impl.h"

SNF_YPLAN(r256, YFFT<8, 8, 4>)
SNF_YPLAN(direct1, YDirect<1>)
SNF_YPLAN(direct2, YDirect<2>)
