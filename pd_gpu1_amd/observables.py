"""Physical observables of a run, computed from the engine's fields.

crack_tip: the crack-tip tracker of the crack-growth driver
(Solve_Plate_Crack_Grow.cu:150-160, commented out in the reference): on the
centre row of an nx x ny plate, the first particle from the left edge whose
damage exceeds 0.38 marks the tip of the left crack branch; its distance
from the plate centre is the current crack half-length.  (The reference
writes (nx/2 - k - 50)*dx, the half-length less the 50 cells of its initial
half-crack at dx = 0.1 mm, i.e. the extension; subtract the initial
half-length from crack_half_length for that.)  Sampling it every few steps
gives the crack speed.

kw_crack_angles: the crack direction of each notch of the Kalthoff-Winkler
specimen (notches cut along x = -rad and x = +rad for y >= 0,
Solve_Kalthoff_Winkler.cu:65-87, tips at (+-rad, 0)), as the angle between
the notch axis (pointing from the impacted edge to the tip, -y) and the
principal direction of the damaged particles below the tip, measured away
from the specimen's centre line.
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


def crack_angle_from_tip(px, py, tip_x, tip_y, outward):
    """Angle in degrees between the axis (0, -1) and the principal direction of
    the points (px, py) about the tip (second moment about the tip, not the
    centroid, so the line passes through the tip); positive when the crack
    turns towards `outward` (+1: +x, -1: -x).  NaN without points."""
    dx, dy = np.asarray(px) - tip_x, np.asarray(py) - tip_y
    if len(dx) < 3:
        return float("nan")
    m = np.array([[np.dot(dx, dx), np.dot(dx, dy)], [np.dot(dx, dy), np.dot(dy, dy)]])
    w, v = np.linalg.eigh(m)
    d = v[:, int(np.argmax(w))]
    if d[1] > 0:
        d = -d  # point away from the notch, into y < tip_y
    ang = np.degrees(np.arccos(np.clip(-d[1], -1.0, 1.0)))
    return float(ang if d[0] * outward >= 0 else -ang)


def kw_crack_angles(x, y, dmg, rad, dx, threshold=0.38, reach=0.04):
    """Crack angles (left notch, right notch) of a Kalthoff-Winkler run: the
    damaged particles (dmg > threshold) within `reach` of a notch tip, at
    least 2 dx below it and on its side of the centre line, all thickness
    layers together."""
    x, y, dmg = np.asarray(x), np.asarray(y), np.asarray(dmg)
    out = []
    for side in (-1.0, 1.0):
        tx = side * rad
        sel = ((dmg > threshold) & (y < -2.0 * dx) & (side * x > 0.0)
               & ((x - tx) ** 2 + y ** 2 <= reach * reach))
        out.append(crack_angle_from_tip(x[sel], y[sel], tx, 0.0, side))
    return tuple(out)
