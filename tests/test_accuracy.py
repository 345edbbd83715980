"""moses_2pdf_amd.accuracy on hand-made states (CPU)."""
import numpy as np

from moses_2pdf_amd.accuracy import depth_error
from moses_2pdf_amd import scenarios as S


def state(h_solid, h_fluid):
    ny, nx = h_solid.shape
    u = np.zeros((7, ny + 6, nx + 6))
    u[0, 3:-3, 3:-3] = h_solid
    u[3, 3:-3, 3:-3] = h_fluid
    u[0, 0, 0] = 1e9   # ghost cells are not counted
    return u


def test_identical_states_have_zero_error():
    h = np.arange(12.0).reshape(3, 4)
    e = depth_error(state(h, h), state(h, h))
    assert e["l1"] == 0 and e["linf"] == 0 and e["runout_iou"] == 1.0


def test_known_l1_linf_and_runout():
    ref = np.array([[0.0, 1.0], [2.0, 1.0]])   # flow depth (scaled) 0, 1, 2, 1 -> sum 4, max 2
    got = np.array([[0.5, 1.0], [1.0, 0.0]])   # |d| = 0.5, 0, 1, 1 -> sum 2.5, max 1
    e = depth_error(state(got, 0 * got), state(ref, 0 * ref), runout_m=0.06)  # threshold 0.6
    assert e["l1"] == 2.5 / 4 and e["linf"] == 1 / 2
    # runout (cells in row-major order): got {1, 2}, ref {1, 2, 3}: intersection 2, union 3
    assert e["runout_cells"] == 2 and e["runout_cells_fp64"] == 3
    assert abs(e["runout_iou"] - 2 / 3) < 1e-15


def test_latin_hypercube_strata_and_ranges():
    x = S.latin_hypercube(64, 3, [(15.0, 30.0, "lin"), (1e5, 5e6, "log")])
    assert x.shape == (64, 2)
    # one sample per stratum in each variable
    u = (x[:, 0] - 15.0) / 15.0
    assert sorted(np.floor(u * 64).astype(int)) == list(range(64))
    v = np.log(x[:, 1] / 1e5) / np.log(50.0)
    assert sorted(np.floor(v * 64).astype(int)) == list(range(64))


def test_ensemble_population_weights_and_relief():
    ms = S.ensemble_members(256)
    from collections import Counter
    c = Counter(m.n for m in ms)
    assert set(c) <= {128, 256, 512, 1024} and sum(c.values()) == 256
    assert c[128] > c[256] > c[512] > c[1024]
    assert abs(S.self_affine_relief(2048) - 1500.0) < 1e-9
    assert abs(S.self_affine_relief(1024) - 1500.0 * 0.5 ** 0.8) < 1e-9
