// Spectral y plans: line lengths 97 (Rader) and 192.
#include "snf_yplan_impl.h"

SNF_YPLAN(rader97, YRader<97, 5, 8, 4, 3>)
SNF_YPLAN(r192, YFFT<8, 8, 3>)
