"""Running DFT of the NTFF contour against the time-domain accumulators.

The time-domain binning is linear in the contour samples, and its bins and
weights are fixed per (angle, cell): a sample taken at base time tb lands on
bins tb + d + (-1, 0, 1) with weights (b, a - b, -a).  The DFT of an
accumulator row at w, sum_m U[m] e^{i w m}, is therefore
sum_cells V(w) e^{i w (d - 1)} (b + (a - b) e^{i w} - a e^{2 i w}) with
V(w) = sum_t sample_t e^{i w tb_t}.  The running DFT keeps V on the device
and projects it once; here the same transform is taken on the host from the
binned accumulators.  The two paths share only the samples, and their
agreement is an identity, so the tolerance is FP64 rounding (1e-10), for the
FP64 path and for the FP32 path (whose samples are combined in FP64 for both).

The run stops a few steps before stepNum so that no bin spills past the end of
its [360][arraySize] row into the next angle (the reference's flat layout,
which the binning keeps); the first bins of a row would take the previous
row's spill, but the contour is still dark during the first steps.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

C0 = 0.7071  # C_0_S (field.h:70)


def _host_dft(U, w):
    """sum_m U[a][m] e^{i w m} for every row a."""
    return U @ np.exp(1j * w * np.arange(U.shape[1]))


def _sim(m, solver, precision, steps, T=5):
    kw = dict(T=T, R=16) if precision == "fp32" else {}
    return m.Simulator(solver, 1, m.field_info(4000, 4000, 50, 10, 500, 0, steps), ntff_time=True,
                       precision=precision, **kw)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("solver", [4, 5, 2, 3])
def test_running_dft_equals_dft_of_time_domain_accumulators(solver, precision):
    import mpifdtd_amd as m
    steps, h = 600, 50
    lams = [500.0, 430.0, 650.0]
    sim = _sim(m, solver, precision, steps)
    try:
        sim.set_dft(lams)
        sim.calc(steps - 6)
        D = sim.dft()
        U = [sim.ntff_time(a) for a in range(3)]
        S = sim.dft_spectra()
    finally:
        sim.finish()
    assert D.shape == (3, 3, 360) and S.shape == (3, 360)
    for f, lam in enumerate(lams):
        w = C0 * 2 * np.pi / (lam / h)
        for a in range(3):
            ref = _host_dft(U[a], w)
            assert np.abs(ref).max() > 0
            assert rel_l2(D[f, a], ref) <= 1e-10, (lam, a, rel_l2(D[f, a], ref))
    # the spectra are the time translation of the same transforms: TM Eth from
    # (Ux, Uy, Wz), TE Eph from (Wx, Wy, Uz), |.|^2 / 8192 (ntffTM.c:161-233)
    assert np.all(np.isfinite(S)) and S.max() > 0
    phi = np.arange(360) * np.pi / 180.0
    ws = C0 * 2 * np.pi / (500.0 / h)  # the source's omega
    coef = np.sqrt(2 * np.pi * C0 / (1j * ws))
    if solver in (2, 3):
        coef /= 4 * np.pi * C0
    z0 = 1.41422712488  # Z_0_S (field.h:75)
    rot = -D[:, 0] * np.sin(phi) + D[:, 1] * np.cos(phi)
    e = coef * (z0 * D[:, 2] - rot) if solver in (2, 4) else coef * (-z0 * rot - D[:, 2])
    assert rel_l2(S, np.abs(e) ** 2 / 8192) <= 1e-12


def test_running_dft_fp32_matches_fp64():
    """The FP32 fields' running DFT agrees with the FP64 fields' to the FP32
    far-field bound, at several wavelengths of a pulse run."""
    import mpifdtd_amd as m
    lams = [450.0, 500.0, 600.0]
    out = {}
    for prec in ("fp64", "fp32"):
        sim = _sim(m, 2, prec, 700)
        try:
            sim.set_dft(lams)
            sim.calc(700)
            out[prec] = sim.dft()
        finally:
            sim.finish()
    for f in range(len(lams)):
        for a in range(3):
            assert rel_l2(out["fp32"][f, a], out["fp64"][f, a]) <= 1e-4, (f, a)


def test_running_dft_block_depth_invariant():
    """FP32: the running DFT does not depend on the temporal block depth."""
    import mpifdtd_amd as m
    res = []
    for T in (1, 8):
        sim = _sim(m, 4, "fp32", 500, T=T)
        try:
            sim.set_dft([500.0, 520.0])
            sim.calc(500)
            res.append(sim.dft())
        finally:
            sim.finish()
    assert rel_l2(res[1], res[0]) <= 1e-12


def test_contour_samples_independent_of_launch_shape():
    """FP32: the per-step contour samples come from the merged, split or two-stream
    tile launches alike: the running DFT (per-cell sums, no atomics) is bit for bit
    the same, and the binned accumulators agree to their atomic summation order."""
    import mpifdtd_amd as m
    D, U = [], []
    for shape in (1, 2, 3, 4, 5, 6, 7, 8, 9):
        sim = m.Simulator(4, 1, m.field_info(6000, 6000, 20, 16, 500, 0, 300), ntff_time=True, precision="fp32",
                          T=8, R=16)
        try:
            sim.set_launch_mode(shape)
            sim.set_dft([500.0, 600.0])
            sim.calc(296)
            D.append(sim.dft())
            U.append([sim.ntff_time(a) for a in range(3)])
        finally:
            sim.finish()
    for k in range(1, len(D)):
        assert np.array_equal(D[k], D[0]), k
        for a in range(3):
            assert rel_l2(U[k][a], U[0][a]) <= 1e-12, (k, a)


def _dft_run(m, solver, precision, steps, lams, ntff_time, T=5, stops=(), reset_at=None):
    kw = dict(T=T, R=16) if precision == "fp32" else {}
    sim = m.Simulator(solver, 1, m.field_info(4000, 4000, 50, 10, 500, 0, steps), ntff_time=ntff_time,
                      precision=precision, **kw)
    try:
        sim.set_dft(lams)
        done = 0
        for s in stops:  # read the DFT part way (the staged steps go in first)
            sim.calc(s - done)
            done = s
            sim.dft()
        if reset_at is not None:
            sim.calc(reset_at - done)
            sim.reset()
            done = 0
        sim.calc(steps - done)
        return sim.dft()
    finally:
        sim.finish()


@pytest.mark.parametrize("nlam", [1, 3, 130])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("solver", [2, 5])
def test_running_dft_staged_equals_per_block(solver, precision, nlam):
    """Without the time-domain binning the samples of up to 128 steps are staged
    and transformed in one launch; with it every block (FP32) or step (FP64) is
    transformed as it is binned.  Same samples, same sums in another order:
    FP64 rounding apart.  130 wavelengths take two frequency blocks of the
    transform, 1 and 3 the narrow forms."""
    import mpifdtd_amd as m
    lams = list(np.linspace(420.0, 700.0, nlam)) if nlam > 1 else [500.0]
    steps = 400
    a = _dft_run(m, solver, precision, steps, lams, ntff_time=False)
    b = _dft_run(m, solver, precision, steps, lams, ntff_time=True)
    assert np.linalg.norm(b) > 0
    assert rel_l2(a, b) <= 1e-12, rel_l2(a, b)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_running_dft_read_part_way_and_reset(precision):
    """Reading the DFT part way (at steps that end inside a staged run) leaves the
    result unchanged; after a reset the DFT starts again from zero."""
    import mpifdtd_amd as m
    lams = [450.0, 500.0, 610.0]
    base = _dft_run(m, 4, precision, 400, lams, ntff_time=False)
    mid = _dft_run(m, 4, precision, 400, lams, ntff_time=False, stops=(55, 200, 330))
    assert np.linalg.norm(base) > 0
    assert rel_l2(mid, base) <= 1e-12, rel_l2(mid, base)
    again = _dft_run(m, 4, precision, 400, lams, ntff_time=False, stops=(60,), reset_at=150)
    assert rel_l2(again, base) <= 1e-12, rel_l2(again, base)
