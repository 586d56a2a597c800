"""The event catalogue (cycle.events): onsets by log-linear interpolation
at the upward threshold crossing, peaks, ends, and merging of crossings
closer than the merge gap, on records whose answers are known in closed
form."""
import numpy as np

from oetqf_amd.cycle import events, recurrence


def _record():
    # two ruptures: v jumps 1e-6 -> 1e-1 between t = 10 and 11 (log-linear
    # crossing of 1e-3 at 10.6), stays above until t = 14, then a second
    # rupture crossing between t = 30 (1e-4) and 31 (1e-2) at 30.5
    t = np.arange(0.0, 40.0)
    v = np.full(t.shape, 1e-6)
    v[11:15] = [1e-1, 0.5, 2e-1, 2e-3]
    v[30] = 1e-4
    v[31:33] = [1e-2, 3.0]
    return t, v


def test_onsets_peaks_and_ends():
    t, v = _record()
    ev = events(t, v, 1e-3)
    assert len(ev) == 2
    assert abs(ev[0]["onset"] - (10.0 + 3.0 / 5.0)) < 1e-12
    assert ev[0]["peak_v"] == 0.5 and ev[0]["t_peak"] == 12.0 and ev[0]["end"] == 14.0
    assert abs(ev[1]["onset"] - 30.5) < 1e-12
    assert ev[1]["peak_v"] == 3.0 and ev[1]["end"] == 32.0
    assert np.allclose(recurrence(ev), [30.5 - 10.6])


def test_merge_gap_joins_close_crossings():
    t, v = _record()
    v[16] = 5e-3  # a flicker above the threshold 2 time units after the first event ends
    assert len(events(t, v, 1e-3)) == 3
    ev = events(t, v, 1e-3, merge_gap=2.5)
    assert len(ev) == 2 and ev[0]["end"] == 16.0 and ev[0]["peak_v"] == 0.5


def test_edges():
    assert events(np.array([0.0]), np.array([1.0]), 1e-3) == []
    t = np.arange(5.0)
    assert events(t, np.full(5, 1.0), 1e-3) == []          # above from the start: no upward crossing
    ev = events(t, np.array([1e-6, 1e-6, 1e-6, 1e-6, 1.0]), 1e-3)  # still above at the end
    assert len(ev) == 1 and ev[0]["end"] == 4.0
