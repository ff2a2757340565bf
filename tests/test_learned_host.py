"""Host logic of the learned step size (gckpp_amd.learned): the labels
derived from logged step histories and the cost classes, on hand-built
histories whose answers follow from the step-size control rules
(FacMin 0.2, FacMax 6, FacSafe 0.9, error exponent 1/3)."""
import numpy as np


def _log(rows):
    """hlog [2*nlog][ncell] from per-cell lists of (H, err)."""
    nlog = max(len(r) for r in rows)
    h = np.zeros((2 * nlog, len(rows)), dtype=np.float32)
    for c, r in enumerate(rows):
        for k, (H, e) in enumerate(r):
            h[2 * k, c] = H
            h[2 * k + 1, c] = e
    return h


def test_labels_follow_the_controller_rules():
    from gckpp_amd import learned
    dt = 1200.0
    rows = [
        # first attempt rejected (err > 1), then accepted at 2e-6: label = the accepted step
        [(1e-5, 40.0), (2e-6, 0.5), (5e-6, 0.3)],
        # accepted and error-limited from the start: the controller's proposal
        # H * FacSafe / err^(1/3) (err = 0.2 > (0.9/6)^3)
        [(1e-5, 0.2), (1.5e-5, 0.3)],
        # never error-limited within the log: last H times FacMax
        [(1e-5, 1e-9), (6e-5, 1e-9), (3.6e-4, 1e-9)],
        # growth capped by dt
        [(100.0, 1e-12), (600.0, 1e-12)],
    ]
    nstep = np.array([3, 2, 3, 2])
    lnh = learned.labels(_log(rows), nstep, dt)
    h = np.exp(lnh)
    assert np.isclose(h[0], 2e-6, rtol=1e-6)
    assert np.isclose(h[1], 1e-5 * 0.9 / 0.2 ** (1.0 / 3.0), rtol=1e-6)
    assert np.isclose(h[2], 3.6e-4 * 6.0, rtol=1e-6)
    assert np.isclose(h[3], dt, rtol=1e-12)


def test_labels_ignore_entries_beyond_the_attempt_count():
    """A cell that stopped after one attempt: later (stale) log entries of
    the array are not read."""
    from gckpp_amd import learned
    rows = [[(1e-5, 1e-9), (1e-3, 0.5)]]
    lnh = learned.labels(_log(rows), np.array([1]), 1200.0)
    assert np.isclose(np.exp(lnh[0]), 1e-5 * 6.0, rtol=1e-6)


def test_classes_are_attempt_quartiles():
    from gckpp_amd import learned
    nstep = np.arange(1, 101)
    thr = np.quantile(nstep, [0.25, 0.5, 0.75])
    cls = learned.classes(nstep, thr)
    assert cls.dtype == np.int32
    assert set(np.unique(cls)) == {0, 1, 2, 3}
    assert (np.diff(cls) >= 0).all()
    assert abs(np.bincount(cls) - 25).max() <= 1
