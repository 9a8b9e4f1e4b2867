"""Writes tests/golden/config2_192x193x192.npz: the oracle's 192x193x192 run
(Lx:Ly:Lz = 2:1:1, S = 1, Re_z = 4000, CFL 0.5, explicit viscous terms, seeded
k^4 exp(-2 (k/k0)^2) field, seed 20260001) after 10 RK3 steps, and
config2_192x193x192_implicit.npz: the same run with the implicit viscous
stages (explicit = 0).

The fields are 25 MB each, so the fixture keeps, per field, its L2 norm and a
sketch S x: 64 seeded complex Gaussian projections (rows regenerated from
SKETCH_SEED by ``sketch``), from which the GPU test estimates the relative L2
error ||S (x - x_ref)|| / ||S x_ref||.  Also kept: the UNER records, time and
offsets.  Run from the repository root: python -m tests.golden.make_config2
"""
import math
import os

import numpy as np

GRID = (192, 193, 192)
RE = 4000.0
NSTEP = 10
SEED = 20260001
SKETCH_SEED = 777
NROW = 64
HERE = os.path.dirname(os.path.abspath(__file__))


def fixture_path(explicit):
    return os.path.join(HERE, "config2_192x193x192%s.npz" % ("" if explicit else "_implicit"))


def sketch(x, seed=SKETCH_SEED, nrow=NROW):
    """nrow complex Gaussian projections of the flattened array x"""
    x = np.ravel(np.asarray(x, dtype=np.complex128))
    rng = np.random.default_rng(seed)
    out = np.empty(nrow, dtype=np.complex128)
    for r in range(nrow):
        g = rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)
        out[r] = np.dot(g, x)
    return out


def write(explicit):
    from oracle.cross import Cross
    from oracle.grid import Grid
    from shear_nnf_amd.ic import make_ic
    vor, phi, u00, w00 = make_ic(*GRID, math.pi, 2 * math.pi, 1.0, seed=SEED)
    g = Grid(*GRID, Re=RE, alp=math.pi, gam=2 * math.pi, Ly=1.0, s=1.0, CFL=0.5)
    c = Cross(g, explicit=explicit, nhist=1)
    c.set_state(vor, phi, u00, w00)
    assert c.run(NSTEP) == NSTEP
    data = {"history": np.array(c.history), "time": c.time, "xoffb": c.xoffb, "xofft": c.xofft,
            "u00": c.u00, "w00": c.w00}
    for nm in ("vor", "phi"):
        f = getattr(c, nm)
        data[nm + "_norm"] = np.linalg.norm(f)
        data[nm + "_sketch"] = sketch(f)
    out = fixture_path(explicit)
    np.savez(out, **data)
    print("wrote", out)


def main():
    for explicit in (1, 0):
        write(explicit)


if __name__ == "__main__":
    main()
