"""Test infrastructure (the oracle): the reference's per-step output
decisions restated on the host, to check the device's event list and
snapshot steps.  Only tests import this.

main_LH.f90:1339-1407 -- after every accepted step k at time x:
  * not slipping and max|V| > velth: an event starts (outfield, slipping,
    dtout = dtout_co, tout = x + dtout, eventcount + 1, moment0 = mean slip,
    hypoloc = maxloc(|V|) in cluster order, onset_time = x);
  * slipping and max|V| < velth / 2: it ends (outfield, not slipping,
    dtout = dtout_inter, tout = x + dtout, moment = mean slip - moment0, row
    (eventcount, k, onset_time, (log10(moment rigid sg) + 5.9) / 1.5, hypoloc),
    coseismic slip |slip - slip at onset|);
  * mod(k, interval) == 0: outfield;
  * x > tout: outfield, tout = x + dtout.
Initial state (main_LH.f90:1188-1195): dtout = dtout_inter, tout = dtout,
not slipping, eventcount 0.
"""
from __future__ import annotations

import math

RIGID = 32.04   # m_const.f90:6


def decisions(steps, velth, dtout_inter, dtout_co, interval, area_sum, k0=0):
    """steps: iterable of (t, vmax, meanslip, argmax0) after each accepted
    step (argmax0: 0-based cluster-order index of the first max |V|).
    Returns (snapshot step numbers, events as dicts, onset step of each event)."""
    slipping = False
    eventcount = 0
    dtout = dtout_inter
    tout = dtout
    moment0 = onset_time = 0.0
    hypoloc = 0
    onset_k = None
    snaps, events, onsets = [], [], []
    for j, (t, vmax, meanslip, loc) in enumerate(steps):
        k = k0 + j + 1
        out = False
        if not slipping and vmax > velth:
            out = True
            slipping = True
            dtout = dtout_co
            tout = t + dtout
            eventcount += 1
            moment0 = meanslip
            hypoloc = int(loc) + 1
            onset_time = t
            onset_k = k
        if slipping and vmax < 0.5 * velth:
            out = True
            slipping = False
            dtout = dtout_inter
            tout = t + dtout
            moment = meanslip - moment0
            events.append({"count": eventcount, "k": k, "onset_time": onset_time,
                           "mw": (math.log10(moment * RIGID * area_sum) + 5.9) / 1.5, "moment": moment,
                           "hypoloc": hypoloc})
            onsets.append(onset_k)
        if interval > 0 and k % interval == 0:
            out = True
        if t > tout:
            out = True
            tout = t + dtout
        if out:
            snaps.append(k)
    return snaps, events, onsets
