"""Shared test configurations: box parameters and seeded initial fields."""
import math

import numpy as np

from shear_nnf_amd.ic import make_ic

# Lx:Ly:Lz = 2:1:1 with Lz = 1 (alp = pi, gam = 2 pi, Ly = 1), S = 1
ALP, GAM, LY = math.pi, 2.0 * math.pi, 1.0


def box_ic(mgalx, my, mgalz, seed=20260001, uprim=0.1):
    return make_ic(mgalx, my, mgalz, ALP, GAM, LY, seed=seed, uprim=uprim)


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))
