"""Batched instances (mpifdtd_batch.h) against the single-instance FP32 path and
the FP64 oracle.

A batch advances uneven instances (sizes, modes, models, wavelengths and step
counts differ) together: device work queues, per-instance progress records and
a replayed hipGraph.  Each instance must end exactly where its own
single-instance run ends; the per-cell arithmetic is the same kernel body, and
the step scalars are formed on the device in FP64 from the same formulas.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

# solver, model, width_nm, height_nm, h, pml, lambda, angle, steps, params
CASES = [
    (4, 1, 6000, 6000, 20, 16, 500, 0, 400, {}),                                # MPI TM Mie, 300 x 300
    (5, 1, 8400, 7200, 20, 16, 560, 30, 450, {}),                               # MPI TE Mie, 420 x 360, oblique
    (4, 2, 5120, 5120, 20, 12, 610, 0, 300, {"ml_layer_num": 3}),               # MPI TM multilayer, stops early
    (5, 5, 6400, 6400, 20, 16, 450, 0, 420, {"zz_layer_num": 3}),               # MPI TE zigzag
]
# every instance runs long enough for its scattered wave to cross the NTFF
# contour: far fields of 1.5e-2 to 5e-2 (oracle), none of them zero
T, R = 4, 16


def _items(m, B):
    return [B.item(s, mo, w, h, hu, p, lam, ang, st, **pr) for (s, mo, w, h, hu, p, lam, ang, st, pr) in CASES]


def _single(m, case):
    s, mo, w, h, hu, p, lam, ang, st, pr = case
    sim = m.Simulator(s, mo, m.field_info(w, h, hu, p, lam, ang, st), params=m.model_params(**pr), ntff_time=False,
                      precision="fp32", T=T, R=R)
    try:
        sim.calc(st)
        return [sim.state(k) for k in (0, 3, 6)], sim.far_field()
    finally:
        sim.finish()


def test_batch_matches_single_instance_runs():
    import mpifdtd_amd as m
    from mpifdtd_amd import batch as B
    print("batch of", len(CASES), "uneven instances", flush=True)
    with B.Batch(_items(m, B), T=T, R=R) as b:
        b.run()
        done = b.sync()
        assert done == [c[8] for c in CASES]
        got = [([b.field(i, k) for k in (0, 3, 6)], b.far_field(i)) for i in range(len(CASES))]
    for i, case in enumerate(CASES):
        f, ff = _single(m, case)
        for k in range(3):
            assert rel_l2(got[i][0][k], f[k]) <= 1e-6, (i, k)
        assert rel_l2(got[i][1], ff) <= 1e-6, i


def test_batch_instance_against_oracle():
    import mpifdtd_amd as m
    from mpifdtd_amd import batch as B
    from oracle import pyoracle as po
    s, mo, w, h, hu, p, lam, ang, st, pr = CASES[1]
    with B.Batch([B.item(s, mo, w, h, hu, p, lam, ang, st, **pr)], T=T, R=R) as b:
        b.run()
        f = [b.field(0, k) for k in (0, 3, 6)]
        ff = b.far_field(0)
    o = po.Sim(width_nm=w, height_nm=h, h_u_nm=hu, pml=p, lambda_nm=lam, angle_deg=ang, stepNum=st, solver=s, model=mo,
               ntff_time=0)
    o.step(st)
    for k, which in enumerate((0, 3, 6)):
        assert rel_l2(f[k], o.field(which)) <= 2e-4, k
    assert rel_l2(ff, o.far_field()) <= 1e-4


def test_batch_chunks_and_graph_replays_agree():
    """Blocks queued in chunks (with syncs between) give the same state as one run
    to completion, which replays the captured graph of 16 blocks."""
    import mpifdtd_amd as m
    from mpifdtd_amd import batch as B
    with B.Batch(_items(m, B), T=T, R=R) as b:
        b.run()
        full = [b.field(i, 0) for i in range(len(CASES))]
    with B.Batch(_items(m, B), T=T, R=R) as b:
        b.run_blocks(5)
        b.sync()
        b.run_blocks(40)
        b.sync()
        b.run()
        part = [b.field(i, 0) for i in range(len(CASES))]
    for a, c in zip(full, part):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("shape", [-1, 8, 9, 10])
def test_batch_rejects_unavailable_launch_shapes(shape):
    """Values outside FAST_LAUNCH_* and the 16-wavefront UPML shapes (single
    grids only) are refused with FDTD_EINVAL, and the batch keeps running in its
    previous shape."""
    import mpifdtd_amd as m
    from mpifdtd_amd import batch as B
    with B.Batch(_items(m, B), T=T, R=R) as b:
        with pytest.raises(RuntimeError, match="launch mode"):
            b.set_launch_mode(shape)
        b.run()
        assert b.sync() == [c[8] for c in CASES]


@pytest.mark.parametrize("shape", [2, 3, 4, 5, 6, 7])
def test_batch_launch_shapes_equal_default(shape):
    """The launches in a row or side by side (a fork and join captured into the
    replayed graph), with either tile kind as 4 or 8 wavefronts, give the same
    state, bit for bit, as the default shape (two streams, 8-wavefront UPML
    tiles); the run is chunked so that eager blocks and graph replays both occur."""
    import mpifdtd_amd as m
    from mpifdtd_amd import batch as B
    with B.Batch(_items(m, B), T=T, R=R) as b:
        b.run()
        ref = [[b.field(i, k) for k in (0, 3, 6)] for i in range(len(CASES))]
    with B.Batch(_items(m, B), T=T, R=R) as b:
        b.set_launch_mode(shape)
        b.run_blocks(5)
        b.sync()
        b.run_blocks(20)
        b.sync()
        b.run()
        assert b.sync() == [c[8] for c in CASES]
        got = [[b.field(i, k) for k in (0, 3, 6)] for i in range(len(CASES))]
    for i in range(len(CASES)):
        for x, y in zip(got[i], ref[i]):
            assert np.array_equal(x, y), i


def test_reference_sweep_runs_as_a_batch():
    """The first structures of the reference's zigzag sweep (fitted field sizes, two
    angles each) run as one batch; every instance reaches its step count and has a
    finite, non-zero far field."""
    import mpifdtd_amd as m
    from mpifdtd_amd import batch as B
    items = B.reference_sweep(5, 5, h_u_nm=20, pml=10, lambda_nm=500, steps=120, start_deg=0, end_deg=30,
                              delta_deg=30)[:6]
    assert len(items) == 6 and len({(it.field.width_nm, it.params.zz_degree) for it in items}) == 3
    with B.Batch(items, T=T, R=R) as b:
        b.run()
        assert b.sync() == [120] * 6
        for i in range(6):
            ff = b.far_field(i)
            assert np.all(np.isfinite(ff)) and np.abs(ff).max() > 0, i
