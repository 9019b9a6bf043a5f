"""Host-side helpers of the map precision sweep (erlbm_amd.maps), CPU only:
the re-run bands are checked against brute force on small synthetic maps."""
import numpy as np

from erlbm_amd.maps import _boundary_distance, _margin_band


def test_boundary_distance_is_chessboard_distance_to_a_class_change():
    rng = np.random.default_rng(4)
    u = (rng.uniform(size=(23, 17)) < 0.15).astype(np.int32)
    u[8:15, 5:12] = 1
    d = _boundary_distance(u, 40)
    # boundary cells: a 4-neighbour of the other class
    b = np.zeros(u.shape, bool)
    b[1:, :] |= u[1:, :] != u[:-1, :]
    b[:-1, :] |= u[:-1, :] != u[1:, :]
    b[:, 1:] |= u[:, 1:] != u[:, :-1]
    b[:, :-1] |= u[:, :-1] != u[:, 1:]
    by, bx = np.nonzero(b)
    for y in range(u.shape[0]):
        for x in range(u.shape[1]):
            want = int(np.min(np.maximum(np.abs(by - y), np.abs(bx - x))))
            assert d[y, x] == min(want, 41), (y, x)


def test_margin_band_is_the_smallest_band_reaching_the_target():
    rng = np.random.default_rng(9)
    n = 4000
    u64 = (rng.uniform(size=n) < 0.4).astype(np.int32)
    margin = np.where(u64 == 1, rng.uniform(2.0, 50.0, n), rng.uniform(0.2, 1.99, n))
    u = u64.copy()
    flip = rng.choice(n, 60, replace=False)       # disagreeing points, margins near the threshold
    u[flip] ^= 1
    margin[flip] = 2.0 * np.exp(rng.uniform(-0.3, 0.3, flip.size))
    margin[flip[:3]] = np.nan                     # non-finite margins go last
    steps = rng.integers(100, 1000, n)
    got = _margin_band({"unstable": u, "margin": margin}, u64, steps, 10.0, 5.0, 2.0)
    with np.errstate(invalid="ignore"):
        c = np.abs(np.log(margin / 2.0))
    c = np.where(np.isfinite(c), c, np.inf)
    best = None
    for delta in np.unique(c):
        band = c <= delta
        if np.mean(np.where(band, u64, u) == u64) >= 0.999:
            best = (delta, int(band.sum()), float(steps[band].sum() / steps.sum() * 10.0))
            break
    assert best is not None
    assert got["points"] == best[1] and abs(got["delta"] - round(best[0], 4)) < 1e-12
    assert abs(got["rerun_seconds"] - round(best[2], 3)) < 1e-9
    assert got["agreement"] >= 0.999 and got["differing"] == 60
