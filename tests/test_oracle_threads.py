"""The oracle's thread-parallel build (row loops of each time-step pass
split over OpenMP threads, used to write the long golden runs) computes the
same bits as the serial build the other tests pin."""

import numpy as np

from media import loam


def test_row_threads_give_the_serial_bits(fdtd, oracle):
    import ctypes as C

    two = [fdtd.Layer(1.3, 420, 190, 1700, 20, 35), fdtd.Layer(0, 700, 320, 2000, 60, 100)]
    models = [
        fdtd.Model(Stack=two, GslsMechanisms=3, MaxRange=7.3, Depth=5.1, Spacing=0.1, PMLCells=12),
        fdtd.Model(Stack=loam(), Relax=fdtd.Fit(25, 30), MaxRange=6, Depth=6, Spacing=0.15, DominantFreq=30),
    ]
    ranges = [0.5, 1.0, 2.5, 4.0]
    serial = []
    for m in models:
        dt = oracle.OracleSim(m).Dt()
        serial.append(oracle.run(m, oracle.ricker(1000, 35, dt, 400), ranges, 400))
    saved = oracle._lib
    try:
        oracle.use_row_threads()
        C.CDLL("libgomp.so.1").omp_set_num_threads(4)
        for m, want in zip(models, serial):
            dt = oracle.OracleSim(m).Dt()
            got = oracle.run(m, oracle.ricker(1000, 35, dt, 400), ranges, 400)
            np.testing.assert_array_equal(got.traces, want.traces)
            assert got.peak == want.peak
    finally:
        oracle._lib = saved


def test_stepwise_run_with_a_checkpoint_equals_run(fdtd, oracle, tmp_path):
    """The golden runs' loop (Sim.Step, the receivers, MaxSpeed every 256
    steps) interrupted by a checkpoint written, the state scrambled by more
    steps, and the checkpoint read back gives Run's bits."""
    two = [fdtd.Layer(1.3, 420, 190, 1700, 20, 35), fdtd.Layer(0, 700, 320, 2000, 60, 100)]
    m = fdtd.Model(Stack=two, GslsMechanisms=3, MaxRange=7.3, Depth=5.1, Spacing=0.1, PMLCells=12)
    o = oracle.OracleSim(m)
    dt = o.Dt()
    steps = 600
    force = oracle.ricker(1000, 35, dt, steps)
    ranges = [0.5, 2.5, 4.0]
    want = oracle.run(m, force, ranges, steps)
    o.Drive(force)
    cols = [o.Column(r)[0] for r in ranges]
    got = np.zeros((len(ranges), steps))
    path = str(tmp_path / "state")
    n = 0
    while n < steps:
        o.Step()
        got[:, n] = [o.SurfaceVelocity(c) for c in cols]
        if n == 299:
            o.checkpoint(path, save=True)
            for _ in range(7):
                o.Step()
            o.checkpoint(path, save=False)
        n += 1
    np.testing.assert_array_equal(got, want.traces)
