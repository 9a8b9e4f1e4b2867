"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Floating-point bar: relative L2 <= 1e-12 per field after 10 RK3 steps
(FP64 path), <= 1e-13 for single operator applications.
"""
import math

import numpy as np
import pytest

from tests.boxes import ALP, GAM, LY, box_ic, rel_l2

pytestmark = pytest.mark.gpu

OP_TOL = 1e-13
STEP_TOL = 1e-12


@pytest.fixture(scope="module")
def lib():
    import shear_nnf_amd as snf
    return snf.load_library()


def _box(mgalx=48, my=49, mgalz=48, Re=1000.0, **kw):
    import shear_nnf_amd as snf
    return snf.ShearBox(mgalx, my, mgalz, Re, alp=ALP, gam=GAM, Ly=LY, **kw)


def _oracle(mgalx=48, my=49, mgalz=48, Re=1000.0, explicit=1, **kw):
    from oracle.grid import Grid
    from oracle.cross import Cross
    g = Grid(mgalx, my, mgalz, Re=Re, alp=ALP, gam=GAM, Ly=LY, s=1.0, CFL=0.5)
    return Cross(g, explicit=explicit, **kw)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("my", [33, 49, 97])
def test_derivyc(lib, order, my):
    from oracle.compact import YOps
    rng = np.random.default_rng(100 + my + order)
    nl = 300
    box = _box(48, my, 48)
    u = rng.normal(size=(my, nl)) + 1j * rng.normal(size=(my, nl))
    th = rng.uniform(-math.pi, math.pi, nl)
    shp = np.exp(1j * th)
    shm = np.conj(shp)
    got = box.derivyc(u, shp, shm, order)
    ops = YOps(my, LY / my)
    ref = ops.deriv(u, shp, shm, order)
    assert rel_l2(got, ref) < OP_TOL


@pytest.mark.parametrize("my", [33, 49, 193])
def test_lapcdy(lib, my):
    from oracle.compact import YOps
    rng = np.random.default_rng(7 + my)
    nl = 257
    box = _box(48, my, 48)
    u = rng.normal(size=(my, nl)) + 1j * rng.normal(size=(my, nl))
    shp = np.exp(1j * rng.uniform(-math.pi, math.pi, nl))
    # the smallest k2 of the DNS is min(alp, gam)^2 = pi^2 (kx index 1, kz 0)
    k2 = np.exp(rng.uniform(math.log(math.pi ** 2), math.log(4e5), nl))
    got = box.lapcdy(u, k2, shp, np.conj(shp))
    ops = YOps(my, LY / my)
    ref = ops.helmholtz(u, k2, shp, np.conj(shp))
    # (d_yy - k2) has cond ~ 4/(h^2 k2): up to ~1.5e4 here, so rounding-level
    # differences reach ~1e-13 relative
    assert rel_l2(got, ref) < 1e-12


@pytest.mark.parametrize("grid", [(48, 49, 48), (32, 33, 32), (96, 97, 64)])
def test_fourxz(lib, grid):
    from oracle import fftxz
    mgalx, my, mgalz = grid
    box = _box(mgalx, my, mgalz)
    rng = np.random.default_rng(3)
    npl = 5
    F = rng.normal(size=(npl, box.kxn, box.mz)) + 1j * rng.normal(size=(npl, box.kxn, box.mz))
    got = box.fourxz(F, 1)
    ref = fftxz.to_phys(F, mgalx, mgalz)
    assert rel_l2(got, ref) < OP_TOL
    f = rng.normal(size=(npl, mgalx, mgalz))
    got2 = box.fourxz(f, -1)
    ref2 = fftxz.to_spec(f, box.kxn, box.mz)
    assert rel_l2(got2, ref2) < OP_TOL


def test_hvhg(lib):
    vor, phi, u00, w00 = box_ic(48, 49, 48)
    rng = np.random.default_rng(11)
    shape = vor.shape
    v = 0.1 * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    dv = 0.1 * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    dom = 0.1 * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    box = _box()
    box.set_state(vor, phi, u00, w00)
    o_dvordy, o_hg, o_hv, rf0u, rf0w = box.hvhg(v, dv, dom)
    c = _oracle()
    c.set_state(vor, phi, u00, w00)
    c.hg, c.hv, c.dvordy = v.copy(), dv.copy(), dom.copy()
    c.istep = 5
    c.hvhg(2, 0)
    assert rel_l2(o_dvordy, c.dvordy) < OP_TOL
    assert rel_l2(o_hg, c.hg) < OP_TOL
    assert rel_l2(o_hv, c.hv) < OP_TOL
    assert rel_l2(rf0u, c.rf0u) < OP_TOL
    assert rel_l2(rf0w, c.rf0w) < OP_TOL


def _hvhg_pair(grid, seed=11):
    vor, phi, u00, w00 = box_ic(*grid)
    rng = np.random.default_rng(seed)
    shape = vor.shape
    v = 0.1 * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    dv = 0.1 * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    dom = 0.1 * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    box = _box(*grid)
    box.set_state(vor, phi, u00, w00)
    got = box.hvhg(v, dv, dom)
    c = _oracle(*grid)
    c.set_state(vor, phi, u00, w00)
    c.hg, c.hv, c.dvordy = v.copy(), dv.copy(), dom.copy()
    c.istep = 5
    c.hvhg(2, 0)
    return got, (c.dvordy, c.hg, c.hv, c.rf0u, c.rf0w)


# planes of 64*R points take the wave-FFT kernels (R = 2, 3, 4, 6)
@pytest.mark.parametrize("grid", [(128, 17, 128), (192, 17, 192), (256, 17, 256),
                                  (384, 17, 384), (192, 17, 384), (384, 17, 128), (512, 17, 512),
                                  (512, 17, 192)])
def test_hvhg_wave_kernels(lib, grid):
    got, ref = _hvhg_pair(grid)
    # 512-point planes: longer transforms (9 FFT stages each way) round a
    # little more, ~1.1e-13 against numpy's FFT
    tol = 2e-13 if max(grid) >= 512 else OP_TOL
    for a, b in zip(got, ref):
        assert rel_l2(a, b) < tol


def test_wave_and_generic_xz_paths_agree(lib, monkeypatch):
    grid = (192, 17, 128)
    vor, phi, u00, w00 = box_ic(*grid)
    rng = np.random.default_rng(5)
    shape = vor.shape
    fields = [0.1 * (rng.normal(size=shape) + 1j * rng.normal(size=shape)) for _ in range(3)]
    outs = []
    for path in ("wave", "generic"):
        monkeypatch.setenv("SNF_XZ_PATH", path)
        box = _box(*grid)
        box.set_state(vor, phi, u00, w00)
        outs.append(box.hvhg(*fields))
    # two FFT factorisations of the same transforms: rounding-level agreement
    for a, b in zip(*outs):
        assert rel_l2(a, b) < OP_TOL


def test_cross_wave_kernels_3_steps(lib):
    st, (vor, phi, u00, w00), box, c = _run_pair(1, 3, grid=(192, 33, 128))
    assert rel_l2(vor, c.vor) < STEP_TOL
    assert rel_l2(phi, c.phi) < STEP_TOL
    assert rel_l2(u00, c.u00) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


# explicit substeps on my = 33, 49, 385 take the spectral y kernels (line
# transforms of 3*11, 7*7, 5*7*11 points); other my the banded sweeps
@pytest.mark.parametrize("grid", [(32, 33, 32), (48, 385, 48)])
def test_cross_spectral_y_3_steps(lib, grid):
    st, (vor, phi, u00, w00), box, c = _run_pair(1, 3, grid=grid)
    assert rel_l2(vor, c.vor) < STEP_TOL
    assert rel_l2(phi, c.phi) < STEP_TOL
    assert rel_l2(u00, c.u00) < STEP_TOL
    assert rel_l2(w00, c.w00) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


# iydealiasing=3: compact filter of vor, phi, U, W at every substep; line
# lengths without a spectral plan (145; 513 on the banded substeps) filter
# through the direct DFT
@pytest.mark.parametrize("grid,explicit,path", [((32, 33, 32), 1, "spectral"),
                                                ((48, 49, 48), 1, "spectral"),
                                                ((48, 49, 48), 0, "spectral"),
                                                ((32, 145, 32), 1, "spectral"),
                                                ((32, 193, 32), 1, "spectral"),
                                                ((32, 193, 32), 0, "spectral"),
                                                ((24, 513, 24), 1, "spectral"),
                                                ((24, 513, 24), 1, "banded")])
def test_cross_compact_filter(lib, grid, explicit, path, monkeypatch):
    monkeypatch.setenv("SNF_Y_PATH", path)
    st, (vor, phi, u00, w00), box, c = _run_pair(explicit, 4, grid=grid, ydealias=3)
    assert rel_l2(vor, c.vor) < STEP_TOL
    assert rel_l2(phi, c.phi) < STEP_TOL
    assert rel_l2(u00, c.u00) < STEP_TOL
    assert rel_l2(w00, c.w00) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


def test_spectral_and_banded_y_paths_agree(lib, monkeypatch):
    grid = (48, 49, 48)
    vor, phi, u00, w00 = box_ic(*grid)
    outs = []
    for path in ("spectral", "banded"):
        monkeypatch.setenv("SNF_Y_PATH", path)
        box = _box(*grid, explicit=1)
        box.set_state(vor, phi, u00, w00)
        box.cross(5)
        outs.append(box.get_state())
    for a, b in zip(*outs):
        assert rel_l2(a, b) < STEP_TOL


def _run_pair(explicit, nstep, grid=(48, 49, 48), Re=1000.0, readflag=1, nhist=1, ydealias=0):
    vor, phi, u00, w00 = box_ic(*grid)
    box = _box(*grid, Re=Re, explicit=explicit, readflag=readflag, nhist=nhist, ydealias=ydealias)
    box.set_state(vor, phi, u00, w00)
    st = box.cross(nstep)
    gv = box.get_state()
    c = _oracle(*grid, Re=Re, explicit=explicit, readflag=readflag, nhist=nhist, ydealias=ydealias)
    c.set_state(vor, phi, u00, w00)
    c.run(nstep)
    return st, gv, box, c


@pytest.mark.parametrize("explicit", [1, 0])
def test_cross_10_steps(lib, explicit):
    st, (vor, phi, u00, w00), box, c = _run_pair(explicit, 10)
    assert st.steps_done == 10
    assert rel_l2(vor, c.vor) < STEP_TOL
    assert rel_l2(phi, c.phi) < STEP_TOL
    assert rel_l2(u00, c.u00) < STEP_TOL
    assert rel_l2(w00, c.w00) < STEP_TOL
    assert abs(st.time - c.time) <= 1e-13 * c.time
    assert abs(st.xofft - c.xofft) <= 1e-13 * abs(c.xofft)
    hist = np.array(box.history)
    ref = np.array(c.history)
    assert hist.shape == ref.shape
    np.testing.assert_allclose(hist, ref, rtol=1e-11, atol=1e-14)


def test_cross_generated_ic_first_dt(lib):
    """readflag=0: the first step uses umax >= 1 (cross.f90:1211-1219)"""
    st, (vor, phi, u00, w00), box, c = _run_pair(1, 3, readflag=0)
    assert rel_l2(vor, c.vor) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("grid", [(48, 49, 48), (96, 97, 96), (128, 33, 64)])
def test_jik_roundtrip(lib, grid):
    """the reference's jik order through the device transposes (tiles of 32
    that do not divide my, mz) and back: bit-exact"""
    vor, phi, u00, w00 = box_ic(*grid)
    box = _box(*grid)
    jv = np.ascontiguousarray(vor.transpose(2, 1, 0))
    jp = np.ascontiguousarray(phi.transpose(2, 1, 0))
    box.set_state_jik(jv, jp, u00, w00, time=1.5, xoffb=-0.25, xofft=0.25)
    v2, p2, u2, w2 = box.get_state()
    assert np.array_equal(v2, vor) and np.array_equal(p2, phi)
    a, b, c2, d, t, xb, xt = box.get_state_jik()
    assert np.array_equal(a, jv) and np.array_equal(b, jp)
    assert (t, xb, xt) == (1.5, -0.25, 0.25)


def test_spectra_and_reset(lib):
    """the spectra of the statistics (cross.f90:1970-2016) and the reset after
    a dump (readwrite.f90:1322-1334) against the oracle"""
    grid = (48, 49, 48)
    vor, phi, u00, w00 = box_ic(*grid)
    box = _box(*grid, explicit=1, nhist=1)
    box.set_state(vor, phi, u00, w00)
    c = _oracle(*grid, explicit=1, nhist=1)
    c.set_state(vor, phi, u00, w00)
    for n in (4, 2):
        box.cross(n)
        c.run(n)
        sp, sp2 = box.spectra()
        st, st2, nacum, tdt = box.statistics()
        assert nacum == c.nacum == n
        assert rel_l2(sp, c.spec) < 1e-11 and rel_l2(sp2, c.spec2) < 1e-11
        assert rel_l2(st.T, c.stats) < 1e-11 and rel_l2(st2.T, c.stats2) < 1e-11
        assert tdt == pytest.approx(c.total_dt, rel=1e-12)
        # Parseval: the spectrum of u'^2 summed over (kx, kz) is the plane sum
        assert np.sum(sp[0]) == pytest.approx(np.sum(st[:, 3]), rel=1e-12)
        box.reset_statistics()
        c.reset_statistics()
        sp, _ = box.spectra()
        st, _, nacum, tdt = box.statistics()
        assert nacum == 0 and tdt == 0.0 and not np.any(sp) and not np.any(st)


# line lengths without a radix plan (my <= 144) take the direct-DFT spectral
# y stage: against the oracle's compact solves and against the banded sweeps
@pytest.mark.parametrize("grid", [(32, 32, 32), (48, 40, 48), (64, 65, 64), (32, 101, 32),
                                  (32, 129, 32), (32, 144, 32), (32, 145, 32)])
def test_cross_direct_dft_y_3_steps(lib, grid):
    st, (vor, phi, u00, w00), box, c = _run_pair(1, 3, grid=grid)
    assert rel_l2(vor, c.vor) < STEP_TOL
    assert rel_l2(phi, c.phi) < STEP_TOL
    assert rel_l2(u00, c.u00) < STEP_TOL
    assert rel_l2(w00, c.w00) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("grid", [(64, 65, 64), (96, 96, 96)])
def test_direct_dft_and_banded_y_paths_agree(lib, monkeypatch, grid):
    vor, phi, u00, w00 = box_ic(*grid)
    outs = []
    for path in ("spectral", "banded"):
        monkeypatch.setenv("SNF_Y_PATH", path)
        box = _box(*grid, explicit=1)
        box.set_state(vor, phi, u00, w00)
        box.cross(5)
        outs.append(box.get_state())
    for a, b in zip(*outs):
        assert rel_l2(a, b) < STEP_TOL


# implicit substeps (cross.f90:797-990) on the spectral y stage: the
# Helmholtz solves with k2 + dtr1 as multipliers, radix and direct plans
@pytest.mark.parametrize("grid", [(32, 33, 32), (48, 385, 48), (64, 65, 64), (32, 96, 32)])
def test_cross_spectral_implicit_3_steps(lib, grid):
    st, (vor, phi, u00, w00), box, c = _run_pair(0, 3, grid=grid)
    assert rel_l2(vor, c.vor) < STEP_TOL
    assert rel_l2(phi, c.phi) < STEP_TOL
    assert rel_l2(u00, c.u00) < STEP_TOL
    assert rel_l2(w00, c.w00) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("grid", [(48, 49, 48), (96, 97, 96)])
def test_spectral_and_banded_implicit_agree(lib, monkeypatch, grid):
    vor, phi, u00, w00 = box_ic(*grid)
    outs = []
    for path in ("spectral", "banded"):
        monkeypatch.setenv("SNF_Y_PATH", path)
        box = _box(*grid, explicit=0)
        box.set_state(vor, phi, u00, w00)
        box.cross(5)
        outs.append(box.get_state())
    for a, b in zip(*outs):
        assert rel_l2(a, b) < STEP_TOL


# radix plans of the y transforms for the ensemble's line lengths and the
# powers of two up to 256 (8*4, 4*4*3, 8*8, 13*5, 8*4*3, 8*4*4, 8*8*3, 8*8*4)
@pytest.mark.parametrize("grid,explicit", [((32, 32, 32), 1), ((32, 48, 32), 1), ((32, 64, 32), 1),
                                           ((32, 65, 32), 0), ((32, 96, 32), 1), ((32, 96, 32), 0),
                                           ((32, 128, 32), 1), ((32, 192, 32), 1),
                                           ((32, 192, 32), 0), ((32, 256, 32), 1)])
def test_cross_radix_y_plans_3_steps(lib, grid, explicit):
    st, (vor, phi, u00, w00), box, c = _run_pair(explicit, 3, grid=grid)
    assert rel_l2(vor, c.vor) < STEP_TOL
    assert rel_l2(phi, c.phi) < STEP_TOL
    assert rel_l2(u00, c.u00) < STEP_TOL
    assert rel_l2(w00, c.w00) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("grid", [(64, 192, 64), (48, 256, 48)])
def test_radix_and_banded_y_paths_agree(lib, monkeypatch, grid):
    vor, phi, u00, w00 = box_ic(*grid)
    outs = []
    for path in ("spectral", "banded"):
        monkeypatch.setenv("SNF_Y_PATH", path)
        box = _box(*grid, explicit=1)
        box.set_state(vor, phi, u00, w00)
        box.cross(5)
        outs.append(box.get_state())
    for a, b in zip(*outs):
        assert rel_l2(a, b) < STEP_TOL


# the direct DFT of line lengths without a radix plan on the matrix cores
# (v_mfma_f64_16x16x4f64, the block's four lines transformed together):
# same bar against the oracle as the per-wave sums it can replace
@pytest.mark.parametrize("grid,explicit,ydealias", [((48, 40, 48), 1, 0), ((32, 101, 32), 0, 0),
                                                    ((64, 99, 64), 1, 0), ((32, 129, 32), 1, 3)])
def test_cross_mfma_dft_y_3_steps(lib, monkeypatch, grid, explicit, ydealias):
    monkeypatch.setenv("SNF_Y_DFT", "mfma")
    st, (vor, phi, u00, w00), box, c = _run_pair(explicit, 3, grid=grid, ydealias=ydealias)
    assert rel_l2(vor, c.vor) < STEP_TOL
    assert rel_l2(phi, c.phi) < STEP_TOL
    assert rel_l2(u00, c.u00) < STEP_TOL
    assert rel_l2(w00, c.w00) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


# prime line lengths by Rader's algorithm (a cyclic convolution over the
# radix plan of my - 1: 97 over 8*4*3, 193 over 8*8*3, 257 over 8*8*4),
# explicit, implicit and with the compact filter, against the oracle
@pytest.mark.parametrize("grid,explicit,ydealias", [((32, 97, 32), 1, 0), ((32, 97, 32), 0, 0),
                                                    ((48, 97, 48), 1, 3), ((32, 193, 32), 1, 0),
                                                    ((32, 193, 32), 0, 0), ((32, 257, 32), 1, 0)])
def test_cross_rader_y_plans_3_steps(lib, grid, explicit, ydealias):
    st, (vor, phi, u00, w00), box, c = _run_pair(explicit, 3, grid=grid, ydealias=ydealias)
    assert rel_l2(vor, c.vor) < STEP_TOL
    assert rel_l2(phi, c.phi) < STEP_TOL
    assert rel_l2(u00, c.u00) < STEP_TOL
    assert rel_l2(w00, c.w00) < STEP_TOL
    np.testing.assert_allclose(np.array(box.history), np.array(c.history), rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("grid,precision", [((64, 193, 64), "fp64"), ((96, 97, 96), "fp64"),
                                            ((96, 97, 96), "mixed")])
def test_rader_and_banded_y_paths_agree(lib, monkeypatch, grid, precision):
    """A Rader plan against the banded sweeps over 5 steps (FP64: to the
    parity bar; mixed: the FP32 inter-stage fields of the spectral stage
    against the banded stage's FP64 ones, to the mixed mode's bound)."""
    vor, phi, u00, w00 = box_ic(*grid)
    outs = []
    for path in ("spectral", "banded"):
        monkeypatch.setenv("SNF_Y_PATH", path)
        box = _box(*grid, explicit=1, precision=precision)
        box.set_state(vor, phi, u00, w00)
        box.cross(5)
        outs.append(box.get_state())
    for a, b in zip(*outs):
        assert rel_l2(a, b) < (STEP_TOL if precision == "fp64" else 1e-5)
