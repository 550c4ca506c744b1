"""Non-finite field detection on the GPU (fdtd_solver_checkFinite and the
periodic check of fdtd_solver_setFiniteCheck), with the condition forced from
the host: a NaN written into one cell of Ez and of Dz through
fdtd_solver_setState (the FP64 MPI step rebuilds Ez from Dz every step, the
FP32 interior update carries Ez itself)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_host_forced_nan_is_counted_and_recorded(precision):
    import mpifdtd_amd as m
    print("finite check,", precision, flush=True)
    fi = m.field_info(4000, 4000, 50, 10, 500, 0, 400)
    sim = m.Simulator(m.MPI_TM_UPML_2D, m.MIE_CYLINDER, fi, precision=precision, T=4)
    try:
        sim.calc(20)
        assert sim.check_finite() == [0, 0, 0]
        assert sim.finite_status() is None
        c = sim.NPXa // 2
        for slot in (0, 2):  # Ez, Dz
            a = sim.state(slot)
            a[c, c] = np.nan
            sim.set_state(slot, a)
        assert sim.check_finite() == [1, 0, 0]
        sim.set_finite_check(4)
        sim.calc(8)
        st = sim.finite_status()
        # the first check ran after the 4th step, the update of time 23; by then the
        # NaN has spread from Ez into its H neighbours and back
        assert st is not None and st[0] == 23 and st[1] > 3, st
        counts = sim.check_finite()
        assert counts[0] > 1 and counts[1] > 0 and counts[2] > 0, counts
        sim.reset()
        assert sim.finite_status() is None and sim.check_finite() == [0, 0, 0]
    finally:
        sim.set_finite_check(0)
        sim.finish()
