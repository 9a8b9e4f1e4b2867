// Spectral y plans: line lengths 385, 32 and 33.
#include "snf_yplan_impl.h"

SNF_YPLAN(r385, YFFT<11, 7, 5>)
SNF_YPLAN(r32, YFFT<8, 4>)
SNF_YPLAN(r33, YFFT<11, 3>)
