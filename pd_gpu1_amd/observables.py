"""Physical observables of a run, computed from the engine's fields.

crack_tip: the crack-tip tracker of the crack-growth driver
(Solve_Plate_Crack_Grow.cu:150-162, commented out in the reference): on the
centre row of an nx x ny plate, the first particle from the left edge whose
damage exceeds 0.38 marks the tip of the left crack branch; its distance
from the plate centre is the current crack half-length.  Sampling it every
few steps gives the crack speed.
"""
import numpy as np


def crack_half_length(dmg, nx, ny, dx, threshold=0.38):
    """Half-length of the damaged zone along the centre row (row ny//2 of an
    nx-by-ny plate stored row by row, x fastest), or 0 if no particle of the
    row's left half exceeds the threshold."""
    row = np.asarray(dmg)[(ny // 2) * nx:(ny // 2) * nx + nx]
    hit = np.nonzero(row[: nx // 2] > threshold)[0]
    if len(hit) == 0:
        return 0.0
    return (nx // 2 - int(hit[0])) * dx


def crack_speed(times, half_lengths):
    """Least-squares crack-tip speed of a sampled half-length history."""
    t = np.asarray(times, dtype=np.float64)
    a = np.asarray(half_lengths, dtype=np.float64)
    if len(t) < 2:
        return 0.0
    return float(np.polyfit(t, a, 1)[0])


def track_crack(prob, nx, ny, dx, dt, steps, every=20, threshold=0.38):
    """Run `steps` Verlet steps of a crack-growth problem, sampling the crack
    half-length every `every` steps.  Returns (times, half_lengths)."""
    times, lengths = [], []
    done = 0
    while done < steps:
        n = min(every, steps - done)
        prob.run_verlet(n)
        done += n
        prob.damage()
        times.append(done * dt)
        lengths.append(crack_half_length(prob.get("DMG"), nx, ny, dx, threshold))
    return times, lengths
