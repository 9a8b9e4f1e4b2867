// Spectral y plans: line lengths 513, 48 and 49.
#include "snf_yplan_impl.h"

SNF_YPLAN(r513, YFFT<9, 3, 19>)
SNF_YPLAN(r48, YFFT<4, 4, 3>)
SNF_YPLAN(r49, YFFT<7, 7>)
