"""Hand-computed cases for the driver's per-shot glue restatement
(tests/glue_oracle.py), which in turn checks the engine's driver."""
import numpy as np

import glue_oracle as G

P = dict(dx=10.0, dz=10.0, dt=0.001, fr=20.0, nx=30, nz=20, nt=200, lpad=2, rpad=3, velfx=0.0, delay=10)


def test_window_interior():
    # source at 100 m, receivers 50..150 m: [50 - 20, 150 + 30] = [30, 180] m -> 16 columns
    fx, nxw, cols = G.window(P, 100.0, np.arange(50.0, 151.0, 10.0))
    assert fx == np.float32(30.0) and nxw == 16
    assert list(cols) == list(range(3, 19))


def test_window_replicates_columns_one_and_nx_minus_two():
    # extent [0, 280] m padded to [-20, 310] m: 34 columns; columns left of
    # x = dx use column 1, right of (nx-2) dx use column nx-2 (getVel's
    # x_min_vel / x_max_vel)
    fx, nxw, cols = G.window(P, 20.0, [0.0, 280.0])
    assert fx == np.float32(-20.0) and nxw == 34
    assert list(cols[:5]) == [1, 1, 1, 1, 2]
    assert list(cols[29:]) == [27, 28, 28, 28, 28]


def test_receiver_depths_and_source_cell():
    fx, nxw, _ = G.window(P, 100.0, [60.0, 80.0])  # [60, 100] m padded to [40, 130] m
    igz = G.receiver_gz(P, fx, nxw, [60.0, 80.0], [20.0, 40.0])
    assert fx == np.float32(40.0) and nxw == 10 and igz[2] == 2 and igz[4] == 4 and igz.sum() == 6
    assert G.source_cell(P, fx, 100.0, 30.0) == (6, 3)


def test_reorder_round_trip():
    fx, nxw, _ = G.window(P, 100.0, [60.0, 80.0, 130.0])  # [40, 160] m: 13 columns
    rec = np.random.default_rng(0).standard_normal((nxw, 5)).astype(np.float32)
    csg = G.reorder(P, fx, nxw, [60.0, 80.0, 130.0], 5, rec=rec, to_gather=True)
    assert nxw == 13 and np.array_equal(csg[0], rec[2]) and np.array_equal(csg[2], rec[9])
    back = G.reorder(P, fx, nxw, [60.0, 80.0, 130.0], 5, csg=csg, to_gather=False)
    assert np.array_equal(back[4], rec[4]) and back[5].sum() == 0


def test_to_global_shifts_one_column():
    # window at 50 m, 12 columns: global column ix takes window column
    # ((ix - 1) dx - 50) / dx = ix - 6 when it lies in [4, nxw - 4], rows [4, nz-4)
    m = np.arange(12 * 10, dtype=np.float32).reshape(12, 10) + 1
    d = G.to_global(P, m, 50.0, 0.0, 30, 10)
    for ix in range(30):
        if 10 <= ix <= 14:
            assert np.array_equal(d[ix, 4:6], m[ix - 6, 4:6]) and d[ix, :4].sum() == 0 and d[ix, 6:].sum() == 0
        else:
            assert d[ix].sum() == 0, ix


def test_remove_direct_mute_time():
    v = np.full((30, 20), 2000.0, np.float32)
    csg = np.ones((2, 200), np.float32)
    out = G.remove_direct(P, csg, v, 100.0, 0.0, [100.0, 200.0], [0.0, 0.0])
    # receiver 1 at the source: t = 0 -> it = delay + 10 + 0.5*5/(fr dt) = 10 + 10 + 125
    assert out[0, :145].sum() == 0 and out[0, 145] == 1
    # receiver 2 ten cells away: t = 10 * 10 / 2000 = 0.05 s = 50 samples
    assert out[1, :195].sum() == 0 and out[1, 195] == 1


def test_plane_geometry_columns_and_rows():
    # whole model [0, 290] m with a 0.01 m margin; a column is
    # floor((gx + 0.01) / 10 + 0.1), so 4.99 m is column 0 and 9 m column 1
    isx, cols = G.plane_geometry(P, [45.0, 290.0], [[-5.0, 0.0, 4.99, 9.0], [289.0, 290.01, 290.02]])
    assert list(isx) == [4, 29]
    assert list(cols[0]) == [-1, 0, 0, 1] and list(cols[1]) == [29, 29, -1]
    assert G.plane_rows(P, 20.0, 10.0) == (2, 1)
