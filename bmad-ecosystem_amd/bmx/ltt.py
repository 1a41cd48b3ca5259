"""Long-term tracking turn driver: ltt_run_beam_mode with tracking_method =
'BMAD' (bsim/modules/lt_tracking_mod.f90:1571-1715) over the device bunch.

Each block of turns is one bmx_track_bunch launch (ltt_track_bmad_bunch,
:3186-3221, tracks ele_start -> ele_start once per turn; here the turns of a
block run in the kernel).  After every block the driver counts the live
particles and stops on no live particle (:1686-1689) or when the dead
fraction reaches dead_cutoff (:1692-1695); every `output_every` turns it
records the bunch averages of ltt_calc_bunch_data (:2391-2404): centroid and
covariance of the live particles, from a device reduction (bmx_bunch_moments).
"""
import numpy as np

from . import tracker


def bunch_averages(ctx, bunch):
    """(n_live, centroid[6], covariance[6, 6]) of the live particles."""
    n, s1, s2 = tracker.bunch_moments(ctx, bunch)
    if n == 0:
        return 0, np.zeros(6), np.zeros((6, 6))
    mean = s1 / n
    cov = s2 / n - np.outer(mean, mean)
    return n, mean, cov


def run_beam_mode(lat, bunch, n_turns, output_every=0, dead_cutoff=1.0, turns_per_launch=10, ix_start=0):
    """Track `bunch` for n_turns turns (fewer if it dies).  Returns a list of
    records {turn, n_live, centroid, cov} at turn 0, every output_every turns
    and at the last turn tracked, and the reason the loop ended."""
    ctx = lat.ctx
    n_tot = bunch.n
    records = []

    def record(turn):
        n, mean, cov = bunch_averages(ctx, bunch)
        records.append({"turn": turn, "n_live": n, "centroid": mean, "cov": cov})
        return n

    n_live = record(0)
    turn = 0
    reason = "turns done"
    step = turns_per_launch if output_every <= 0 else min(turns_per_launch, output_every)
    while turn < n_turns:
        k = min(step, n_turns - turn)
        if output_every > 0:
            k = min(k, output_every - turn % output_every)
        tracker.track_bunch(lat, bunch, n_turns=k, ix_start=ix_start, ix_end=ix_start)
        turn += k
        if output_every > 0 and turn % output_every == 0 or turn == n_turns:
            n_live = record(turn)
        else:
            n_live = int((bunch.state == 1).sum().item())
        if n_live == 0:
            reason = "no live particles"
            break
        if n_tot - n_live >= int(round(dead_cutoff * n_tot)):
            reason = "dead cutoff"
            break
    if records[-1]["turn"] != turn:
        record(turn)
    return records, reason
