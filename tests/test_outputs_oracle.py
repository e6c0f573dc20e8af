"""CPU: the host restatement of the reference's output decisions
(oracle/outputs.py) on hand-built step sequences whose snapshot steps and
events are worked out by hand, and its Mw against the Hanks-Kanamori
definition with the units spelled out."""
import math

from oracle import outputs as OUT


def test_event_state_machine_by_hand():
    # k:      1     2     3     4      5     6     7     8     9     10
    # vmax: 1e-3  2e-2  3e-2  4e-3  1e-3  1e-3  2e-2  6e-3  4.9e-3  1e-4
    # t:       1     2   2.5   3.6     4     5     6   6.5    8.1   9
    vm = [1e-3, 2e-2, 3e-2, 4e-3, 1e-3, 1e-3, 2e-2, 6e-3, 4.9e-3, 1e-4]
    ts = [1.0, 2.0, 2.5, 3.6, 4.0, 5.0, 6.0, 6.5, 8.1, 9.0]
    slip = [0.0, 0.1, 0.5, 1.0, 1.0, 1.0, 1.2, 1.6, 2.0, 2.0]
    loc = [0, 7, 3, 3, 0, 0, 11, 2, 2, 0]
    steps = list(zip(ts, vm, slip, loc))
    snaps, events, onsets = OUT.decisions(steps, velth=1e-2, dtout_inter=100.0, dtout_co=1.0, interval=5,
                                          area_sum=2.0)
    # onset at k=2 (tout = 3): k=3 t=2.5 no; k=4 t=3.6 < 5e-3 ends the event
    # (tout = 103); k=5 by interval; onset at k=7 (tout = 7); k=8 t=6.5 no;
    # k=9 t=8.1 > 7 and 4.9e-3 < 5e-3: end (tout = 108.1); k=10 by interval
    assert snaps == [2, 4, 5, 7, 9, 10]
    assert onsets == [2, 7]
    assert [(e["count"], e["k"], e["onset_time"], e["hypoloc"]) for e in events] == [(1, 4, 2.0, 8), (2, 9, 6.0, 12)]
    assert abs(events[0]["moment"] - 0.9) < 1e-15 and abs(events[1]["moment"] - 0.8) < 1e-15


def test_time_outputs_follow_tout():
    # no event: tout starts at dtout_inter and advances by it after each output
    steps = [(t, 1e-9, 0.0, 0) for t in (4.0, 9.0, 10.5, 16.0, 19.0, 30.0)]
    snaps, events, _ = OUT.decisions(steps, velth=1e-2, dtout_inter=5.0, dtout_co=1.0, interval=0, area_sum=1.0)
    # tout 5: t=9 > 5 -> out, tout 14; t=16 -> out, tout 21; t=30 -> out
    assert snaps == [2, 4, 6] and events == []


def test_mw_is_hanks_kanamori():
    # mean slip D = 2 m over A = 1e4 km^2 with mu = 32.04 GPa:
    # M0 = mu A D = 32.04e9 Pa * 1e10 m^2 * 2 m; Mw = 2/3 (log10 M0 - 9.1)
    steps = [(1.0, 1.0, 0.0, 0), (2.0, 1e-4, 2.0, 0)]
    _, events, _ = OUT.decisions(steps, velth=1e-2, dtout_inter=100.0, dtout_co=1.0, interval=0, area_sum=1e4)
    m0 = 32.04e9 * 1e4 * 1e6 * 2.0
    assert abs(events[0]["mw"] - (2.0 / 3.0) * (math.log10(m0) - 9.1)) < 1e-12
