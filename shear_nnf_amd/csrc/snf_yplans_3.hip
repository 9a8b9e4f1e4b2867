// Spectral y plans: line lengths 193 (Rader), 96 and 128.
#include "snf_yplan_impl.h"

SNF_YPLAN(rader193, YRader<193, 5, 8, 8, 3>)
SNF_YPLAN(r96, YFFT<8, 4, 3>)
SNF_YPLAN(r128, YFFT<8, 4, 4>)
