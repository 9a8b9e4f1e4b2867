// Spectral y plans: line lengths 257 (Rader), 64 and 65.
#include "snf_yplan_impl.h"

SNF_YPLAN(rader257, YRader<257, 3, 8, 8, 4>)
SNF_YPLAN(r64, YFFT<8, 8>)
SNF_YPLAN(r65, YFFT<13, 5>)
