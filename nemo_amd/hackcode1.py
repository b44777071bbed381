"""hackcode1 run loop: leapfrog cadence, diagnostics frames and snapshot output.

Mirrors nemo_main/stepsystem/output of the reference
(src/nbody/evolve/hackcode/hackcode1/code.c:57-64,159-191 and
code_io.c:302-345): the loop runs while tnow < tstop + 0.1/freq, each step
does the force pass, then output() (diagnostics after the force pass, a
minor frame with Time+Diagnostics when minor_tout - 0.01/freq <= tnow, a
major frame that adds Mass and PhaseSpace when tout - 0.01/freq <= tnow),
then the drift.  The stepper does the arithmetic: it must offer
force() -> (n2b, nbc), diagnostics() -> 36 doubles, drift(), and the
attributes mass, pos, vel (state at output time after force()).
"""
from __future__ import annotations

from dataclasses import dataclass

from . import snapshot


@dataclass
class LogRow:
    tnow: float
    etot: float
    virial: float
    nttot: int
    nbavg: int
    ncavg: int


def run(stepper, tstop=2.0, freq=32.0, freqout=4.0, minor_freqout=32.0, options="mass,phase",
        history="hackcode1", out: bytearray | None = None):
    """Run to tstop; returns (log rows, output bytes or None)."""
    opts = [o.strip() for o in options.split(",")]
    if out is not None:
        snapshot.put_history(out, history)
    tnow = 0.0
    tout = minor_tout = tnow
    firstmass = True
    rows = []
    n = len(stepper.mass)
    while tnow < tstop + 0.1 / freq:
        n2b, nbc = stepper.force()
        diag = stepper.diagnostics()
        rows.append(LogRow(tnow, float(diag[0]), float(diag[1] / diag[2]), n2b + nbc,
                           int(float(n2b) / float(n)), int(float(nbc) / float(n))))
        time_bit = major = write_mass = False
        if minor_freqout > 0.0 and (minor_tout - 0.01 / freq) <= tnow:
            minor_tout += 1.0 / minor_freqout
            time_bit = True
        if freqout > 0.0 and (tout - 0.01 / freq) <= tnow:
            tout += 1.0 / freqout
            major = True
            if "mass" in opts or firstmass:
                write_mass = True
                firstmass = False
        if (time_bit or major) and out is not None:
            pos = stepper.pos if major else None
            vel = stepper.vel if major else None
            snapshot.put_snapshot(out, stepper.mass, pos, vel, tnow,
                                  write_mass=write_mass, write_phase=major, diag=diag)
        stepper.drift()
        tnow = tnow + 1.0 / freq
    return rows, out
