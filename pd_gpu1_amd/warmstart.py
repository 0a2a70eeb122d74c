"""Learned warm start for quasi-static ADR load stepping (no reference
counterpart; the reference restarts ADR from the previous solution,
Solve_Plate_Tensile.cu:108-131).

A plate with a pre-crack is loaded in displacement increments: the top strip
rows move by +du and the bottom rows by -du, then ADR relaxes the free
particles until the relative residual is below tol.  Plain ADR starts each
increment from the previous solution with only the strips moved.  The warm
start also moves the free particles by a predicted increment:

* while no bond broke in the previous increment, the response is linear in
  the load and the previous increment du_prev is the prediction;
* for the first increment of a plate (no history) and after an increment in
  which bonds broke, the prediction is the uniform-stretch field between the
  strips, du_base = (0, du * y / y_strip), plus the correction the
  message-passing network (pd_mpnn_*) predicts from the plate's bond graph.
  Without the network (the baseline the network is compared with) du_base
  alone is used.  The loaded plate with its centred crack is symmetric under
  the point reflection p -> -p about the plate centre, so while the broken
  bonds are too its solution is odd, u(-p) = -u(p); the network output is
  projected onto that odd part before it is applied.  An even part in the
  guess (a rigid-body rotation, for one) is pure error that ADR damps only
  slowly; removing it is what lets the x correction help instead of hurt.

ADR then converges from that guess to the same tolerance, so accuracy is
unchanged and only the iteration count moves.  Fallback (Plate.step_warm):
by default the configuration's rule, plain ADR for 50 iterations from the
plain start against 50 iterations from the guess, the plain run taking over
when the guess's residual is the larger (probe_then_solve), and a restart
from the plain start when the run from the guess does not reach tol;
optionally a second handle of the same plate runs plain ADR at the same time
and the first of the two runs to reach tol gives the increment.

Node features [N][16]: position relative to the plate centre / half side (2),
du_prev / du (2), damage, top-strip flag, bottom-strip flag, 1, and eight
features of the node against the pre-crack (crack_features).  Edge
features [E][3]: reference bond vector / delta (2), intact flag.  Targets
[N][2]: (du_next - du_base) / du, weighted 1 on free particles and 0 on the
strips (their motion is prescribed).  Training samples are the increments the network is
used for: first increments and increments that follow bond breakage.
"""
import math
import time

import numpy as np

from . import _lib
from . import batch as PB
from .engine import Problem, gen_plate_strips
from .mpnn import Mpnn

DU = 1.0e-6
DX = 0.5e-3
# Factor on the network's correction.  Chosen on validation plates (index
# 128-191 of the evaluation family, disjoint from the plates the evaluation
# reports): the full correction gives the lowest total and median iteration
# count but a heavy tail (2.0x the uniform-stretch start's iterations on one
# plate in 64, worse on 10); 0.75 was worse on 5 of 64, by 1.23x at most,
# and cut the total by 20%.
NET_SCALE = 0.75


def plate_layout(n):
    """Particles of an n x n plate with 3 strip rows above and below (Coord.cu:47-66):
    free particles first, then the top strip rows, then the bottom ones."""
    x, y = gen_plate_strips(n, n, DX)
    n_int = n * n
    top = np.zeros(len(x), bool)
    bot = np.zeros(len(x), bool)
    top[n_int:n_int + 3 * n] = True
    bot[n_int + 3 * n:] = True
    return x, y, n_int, top, bot


def base_increment(x, y, n_int, top):
    """Uniform-stretch increment between the strips: u_y = du * (y - c) / (y_top - c)
    with c the plate centre and y_top the mean height of the top strip rows."""
    cy = y[:n_int].mean()
    ys = y[top].mean() - cy
    b = np.zeros((2, len(x)))
    b[1] = DU * np.clip((y - cy) / ys, -1.0, 1.0)
    return b


def plate_half(x, y, n_int):
    """Half side of the free plate (the unit of the normalized positions)."""
    return 0.5 * max(np.ptp(x[:n_int]), np.ptp(y[:n_int]), DX)


def normalized_positions(x, y, n_int):
    cx, cy = x[:n_int].mean(), y[:n_int].mean()
    half = plate_half(x, y, n_int)
    return np.stack([(x - cx) / half, (y - cy) / half])


def crack_features(pos, clen, angle, half):
    """[8][N] description of each node against the centred pre-crack (the
    global context a few message-passing layers cannot see: the crack's
    perturbation of the displacement field spans its whole length).  With
    (s, t) the node's coordinates along and across the crack in units of the
    crack half-length a: s/3 and t/3 clipped to [-1, 1]; the opening profile
    of a straight crack under remote tension, sign(t) sqrt(1 - s^2) for
    |s| < 1 (0 beyond the tips); a / half side; cos and sin of the crack
    angle; 1 / (1 + r^2) with r the distance to the nearer tip; and
    1 / (1 + s^2 + t^2).  All zero for a plate without a crack."""
    N = pos.shape[1]
    out = np.zeros((8, N))
    a = 0.5 * clen / half
    if not clen > 0.0:
        return out
    c, sn = math.cos(angle), math.sin(angle)
    s = (pos[0] * c + pos[1] * sn) / a
    t = (-pos[0] * sn + pos[1] * c) / a
    out[0] = np.clip(s / 3.0, -1.0, 1.0)
    out[1] = np.clip(t / 3.0, -1.0, 1.0)
    out[2] = np.sign(t) * np.sqrt(np.clip(1.0 - s * s, 0.0, None))
    out[3] = a
    out[4] = c
    out[5] = sn
    r2 = np.minimum((s - 1.0) ** 2, (s + 1.0) ** 2) + t * t
    out[6] = 1.0 / (1.0 + r2)
    out[7] = 1.0 / (1.0 + s * s + t * t)
    return out


def edge_features(x, y, row_ptr, col, mask, delta):
    rows = np.repeat(np.arange(len(row_ptr) - 1), np.diff(row_ptr))
    ef = np.empty((len(col), 3), np.float32)
    ef[:, 0] = (x[col] - x[rows]) / delta
    ef[:, 1] = (y[col] - y[rows]) / delta
    ef[:, 2] = mask
    return ef


def damage(row_ptr, mask):
    """Unweighted broken-bond fraction per particle."""
    deg = np.diff(row_ptr)
    intact = np.add.reduceat(mask.astype(np.float64), row_ptr[:-1]) if len(mask) else np.zeros(len(deg))
    intact = np.where(deg > 0, intact, 0.0)
    return np.where(deg > 0, 1.0 - intact / np.maximum(deg, 1), 0.0)


def node_features(pos, du_prev, dmg, top, bot, crack):
    """pos: [2][N] coordinates relative to the plate centre / half side;
    crack: [8][N] from crack_features."""
    N = pos.shape[1]
    f = np.empty((N, 16), np.float32)
    f[:, 0:2] = pos.T
    f[:, 2:4] = (du_prev / DU).T
    f[:, 4] = dmg
    f[:, 5] = top
    f[:, 6] = bot
    f[:, 7] = 1.0
    f[:, 8:16] = crack.T
    return f


# ------------------------------------------------------------ training data
def generate(specs, n_inc, tol=1e-6, max_iters=400000, block=1024, group=4, log=None):
    """Run the batched ADR on ``specs`` for ``n_inc`` increments and keep every
    converged state.  Returns one record per specimen with its graph and the
    states u_0 = 0, u_1 .. u_n_inc (float64 [2][N]) and intact masks."""
    p = PB.plate_params(DX)
    recs = []
    t0 = time.perf_counter()
    with PB.Batch(p, specs) as b:
        b.set_group(group)
        for s, (n, _, clen, ang) in enumerate(specs):
            rp, col, mask = b.csr(s)
            x, y, n_int, top, bot = plate_layout(n)
            recs.append(dict(n=n, n_int=n_int, rp=rp, col=col, x=x, y=y, top=top, bot=bot, clen=clen, angle=ang,
                             u=[np.zeros((2, len(x)))], mask=[mask]))
        iters = []
        for m in range(n_inc):
            its, res, ms = b.increment(DU, tol, max_iters, block=block)
            iters.append(its)
            for s, r in enumerate(recs):
                r["u"].append(b.get(s, "U"))
                r["mask"].append(b.csr(s)[2])
            if log:
                log(f"[data] increment {m}: iterations median {int(np.median(its))} max {its.max()}, "
                    f"{ms:.0f} ms, max residual {res.max():.2e}")
    return recs, np.array(iters), time.perf_counter() - t0


def uses_network(k, mask_before, mask_prev):
    """The network predicts increment k (0-based) when there is no previous
    increment or when bonds broke during the previous one."""
    return k == 0 or mask_prev is None or not np.array_equal(mask_before, mask_prev)


def samples(rec, delta):
    """(node features, edge features, target, weight) of the increments the
    network predicts (see uses_network)."""
    us, masks = rec["u"], rec["mask"]
    free = np.zeros(len(rec["x"]), np.float32)
    free[: rec["n_int"]] = 1.0
    pos = normalized_positions(rec["x"], rec["y"], rec["n_int"])
    cf = crack_features(pos, rec["clen"], rec["angle"], plate_half(rec["x"], rec["y"], rec["n_int"]))
    base = base_increment(rec["x"], rec["y"], rec["n_int"], rec["top"])
    for k in range(1, len(us)):
        if not uses_network(k - 1, masks[k - 1], masks[k - 2] if k >= 2 else None):
            continue
        prev = us[k - 1] - us[k - 2] if k >= 2 else np.zeros_like(us[0])
        mask = masks[k - 1]
        x = node_features(pos, prev, damage(rec["rp"], mask), rec["top"], rec["bot"], cf)
        ef = edge_features(rec["x"], rec["y"], rec["rp"], rec["col"], mask, delta)
        t = ((us[k] - us[k - 1] - base) / DU).T.astype(np.float32)
        yield x, ef, t, free


class SideGraph:
    """What every plate of side n shares: the lattice with its strips, the
    bond graph (CSR), the normalized positions, the uniform-stretch
    increment and the static edge features (reference bond vector / delta;
    column 2, the intact flag, is filled per sample)."""

    def __init__(self, n, rp, col, delta):
        self.n = n
        self.x, self.y, self.n_int, self.top, self.bot = plate_layout(n)
        self.rp, self.col = rp, col
        self.pos = normalized_positions(self.x, self.y, self.n_int)
        self.half = plate_half(self.x, self.y, self.n_int)
        self.base = base_increment(self.x, self.y, self.n_int, self.top)
        self.ef = edge_features(self.x, self.y, rp, col, np.ones(len(col), np.uint8), delta)
        self.free = np.zeros(len(self.x), np.float32)
        self.free[: self.n_int] = 1.0


def generate_first(specs, sides, delta, tol=1e-6, max_iters=400000, block=1024, group=4, log=None):
    """First increments of ``specs`` by the batched ADR, kept compactly: per
    specimen its side, crack, intact mask before the increment (packed bits)
    and the target (u_1 - du_base) / du as float32 [N][2]; the graph of each
    side goes into ``sides`` (n -> SideGraph) once.  Returns (records,
    iterations, seconds)."""
    p = PB.plate_params(DX)
    recs = []
    t0 = time.perf_counter()
    with PB.Batch(p, specs) as b:
        b.set_group(group)
        masks = []
        for s, (n, _, clen, ang) in enumerate(specs):
            rp, col, mask = b.csr(s)
            if n not in sides:
                sides[n] = SideGraph(n, rp, col, delta)
            masks.append(np.packbits(mask, bitorder="little"))
        its, res, ms = b.increment(DU, tol, max_iters, block=block)
        for s, (n, _, clen, ang) in enumerate(specs):
            u1 = b.get(s, "U")
            t = ((u1 - sides[n].base) / DU).T.astype(np.float32)
            recs.append(dict(n=n, clen=clen, angle=ang, mask=masks[s], nb=len(sides[n].col), t=t))
        if log:
            log(f"[data] {len(specs)} first increments: iterations median {int(np.median(its))} max {its.max()}, "
                f"{ms:.0f} ms, max residual {res.max():.2e}")
    return recs, its, time.perf_counter() - t0


def first_sample(rec, side):
    """(node features, edge features, target, weight) of a compact
    first-increment record; the edge features are the side's array with the
    intact flags of this record written into column 2 (valid until the
    side's next sample)."""
    mask = np.unpackbits(rec["mask"], count=rec["nb"], bitorder="little")
    cf = crack_features(side.pos, rec["clen"], rec["angle"], side.half)
    x = node_features(side.pos, np.zeros((2, len(side.x))), damage(side.rp, mask), side.top, side.bot, cf)
    side.ef[:, 2] = mask
    return x, side.ef, rec["t"], side.free


def epoch_order(recs, rng, window=512):
    """A random permutation of the records, sorted by side within windows of
    `window` consecutive records so that the graph changes less often."""
    perm = rng.permutation(len(recs))
    out = []
    for k in range(0, len(perm), window):
        w = perm[k:k + window]
        out.extend(w[np.argsort([recs[i]["n"] for i in w], kind="stable")])
    return out


def train_first(net, recs, sides, epochs, lr=2e-3, lr_end=5e-5, seed=0, log=None, max_seconds=None, window=512):
    """Adam over compact first-increment records (generate_first).  The
    learning rate decays geometrically from lr to lr_end with the training's
    progress, the larger of the fraction of steps done and, with
    max_seconds, the fraction of that time used; training stops after the
    last step or once max_seconds have passed.  Returns the mean loss of each
    (possibly partial) epoch and the number of steps taken."""
    rng = np.random.default_rng(seed)
    total = epochs * len(recs)
    step = 0
    hist = []
    t0 = time.perf_counter()
    cur = None
    for ep in range(epochs):
        losses = []
        for i in epoch_order(recs, rng, window):
            used = (time.perf_counter() - t0) / max_seconds if max_seconds else 0.0
            if used >= 1.0:
                break
            prog = max(step / max(total - 1, 1), used)
            r = recs[i]
            side = sides[r["n"]]
            if cur is not side:
                net.set_graph(side.rp, side.col)
                cur = side
            x, ef, t, w = first_sample(r, side)
            net.set_features(x, ef, t, w)
            losses.append(net.train_step(lr * (lr_end / lr) ** prog))
            step += 1
            if log and step % 2000 == 0:
                log(f"[train] step {step}: loss (last 2000) {np.mean(losses[-2000:]):.4e}, "
                    f"{time.perf_counter() - t0:.0f} s")
        hist.append(float(np.mean(losses)) if losses else float("nan"))
        if log:
            log(f"[train] epoch {ep}: mean loss {hist[-1]:.4e} over {len(losses)} samples")
        if len(losses) < len(recs):
            break
    return hist, step


def train(net, recs, delta, epochs, lr=1e-3, lr_end=1e-4, seed=0, log=None):
    """Adam over all (specimen, increment) samples; the learning rate decays
    geometrically from lr to lr_end.  Returns the mean loss of each epoch."""
    rng = np.random.default_rng(seed)
    n_samples = sum(1 for r in recs for _ in samples(r, delta))
    total = epochs * n_samples
    step = 0
    hist = []
    for ep in range(epochs):
        losses = []
        for s in rng.permutation(len(recs)):
            r = recs[s]
            net.set_graph(r["rp"], r["col"])
            for x, ef, t, w in samples(r, delta):
                rate = lr * (lr_end / lr) ** (step / max(total - 1, 1))
                net.set_features(x, ef, t, w)
                losses.append(net.train_step(rate))
                step += 1
        hist.append(float(np.mean(losses)))
        if log:
            log(f"[train] epoch {ep}: mean loss {hist[-1]:.4e} over {len(losses)} samples")
    return hist


# --------------------------------------------------------------- evaluation
class Plate:
    """One pre-cracked plate on its own pd_handle, loaded increment by increment.

    Set-up also prepares, once per plate, what the network path reuses every
    time it predicts: the static node and edge features (positions, the
    pre-crack frame, strip flags, reference bond vectors) and the point
    reflection maps (mirror_maps)."""

    def __init__(self, n, crack_len, angle, device=0):
        p = PB.plate_params(DX)
        p.n_top = p.n_bot = 3 * n
        self.x, self.y, self.n_int, self.top, self.bot = plate_layout(n)
        self.prob = Problem(p, device)
        self.prob.set_positions(self.x, self.y, None, n_int=self.n_int)
        self.prob.build_neighbors()
        self.prob.init_state(static=True)
        if crack_len > 0:
            self.prob.set_crack_2d(crack_len, 0.0, 0.0, angle)
        self.delta = p.delta
        self.rp, self.col = self.prob.csr()
        self.pos = normalized_positions(self.x, self.y, self.n_int)
        self.cf = crack_features(self.pos, crack_len, angle, plate_half(self.x, self.y, self.n_int))
        self.base = base_increment(self.x, self.y, self.n_int, self.top)
        self.du_prev = np.zeros((2, len(self.x)))
        self.mask_prev = None  # intact mask before the previous increment
        self.k = 0
        self._static_x = node_features(self.pos, np.zeros((2, len(self.x))), np.zeros(len(self.x)), self.top,
                                       self.bot, self.cf)
        self._ef = edge_features(self.x, self.y, self.rp, self.col, np.ones(len(self.col), np.uint8), self.delta)
        self._mir = None
        self.mirror_maps()

    def close(self):
        self.prob.close()

    def features(self, mask):
        """Node and edge features of the current state (intact mask `mask`)."""
        x = self._static_x.copy()
        x[:, 2:4] = (self.du_prev / DU).T
        x[:, 4] = damage(self.rp, mask)
        self._ef[:, 2] = mask
        return x, self._ef

    def _solve(self, tol, max_iters, probe):
        it1, r1, ok = self.prob.adr_solve(0, min(probe, max_iters), tol)
        if ok or it1 < probe:
            return it1, r1, r1
        it2, r2, _ = self.prob.adr_solve(probe, max_iters - probe, tol)
        return it1 + it2, r2, r1

    def step_plain(self, tol=1e-6, max_iters=400000, probe=50):
        """Plain ADR increment.  Returns (iterations, residual, residual after probe, seconds)."""
        t = time.perf_counter()
        u0 = self.prob.get("U")
        self.prob.strip_displace(DU, -DU)
        its, r, r_probe = self._solve(tol, max_iters, probe)
        self.prob.sync()
        sec = time.perf_counter() - t
        self.du_prev = self.prob.get("U") - u0
        self.k += 1
        return its, r, r_probe, sec

    def mirror_maps(self):
        """Point reflection p -> -p about the plate centre, which maps the
        lattice, the centred crack and the strip loading (top +du, bottom
        -du) onto themselves with the displacement reversed: the particle
        index of each particle's mirror image and the CSR position of each
        bond's mirror bond (computed once per plate)."""
        if self._mir is None:
            cx, cy = self.x[: self.n_int].mean(), self.y[: self.n_int].mean()
            ix = np.rint((self.x - cx) / DX * 2.0).astype(np.int64)
            iy = np.rint((self.y - cy) / DX * 2.0).astype(np.int64)
            M = 4 * (int(max(np.abs(ix).max(), np.abs(iy).max())) + 2)
            key = (ix + M) * (2 * M + 1) + (iy + M)
            order = np.argsort(key)
            mkey = (-ix + M) * (2 * M + 1) + (-iy + M)
            pos = np.searchsorted(key[order], mkey)
            mir = order[np.minimum(pos, len(order) - 1)]
            assert np.array_equal(key[mir], mkey), "the layout is not point-symmetric"
            N = len(self.x)
            rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(self.rp))
            bkey = rows * N + self.col
            bmir = np.searchsorted(bkey, mir[rows] * N + mir[self.col])
            assert np.array_equal(bkey[bmir], mir[rows] * N + mir[self.col])
            self._mir, self._bmir = mir, bmir
        return self._mir, self._bmir

    def step_warm(self, net, tol=1e-6, max_iters=400000, use_net=True, fallback=None, chunk=256, probe=50):
        """Warm-started increment: the previous increment du_prev, or where
        uses_network says so the uniform-stretch increment plus (when use_net)
        the network's correction.  The problem is symmetric under the point
        reflection about the plate centre while the broken bonds are, so its
        solution is odd under it; the network's correction is projected onto
        that odd part (its even part, which is pure error, is removed) and the
        network is used only while the symmetry holds.

        Fallback to plain ADR: when the run from the guess does not reach tol
        within max_iters (or its state goes non-finite), the increment
        restarts from the plain start (previous solution, strips moved) and
        the iterations of both runs are counted.  With `fallback` (a Plate of
        the same specimen, used only as a second solver) plain ADR instead
        runs from the plain start at the same time, on its own handle and
        stream, while this plate runs from the guess, both in chunks of
        `chunk` iterations on two host threads; the first to reach tol gives
        the increment (the other stops at its next chunk boundary) and the
        result is this plate's state.  Without `fallback`, the network's
        guess is checked by the warm-start configuration's rule when `probe`
        > 0 (default 50, probe_then_solve: plain ADR from the plain start and
        ADR from the guess, `probe` iterations each, compared by residual),
        every iteration counted; probe = 0 keeps only the restart.  Returns
        (iterations, residual, seconds, fell_back, predict seconds,
        used_net)."""
        t = time.perf_counter()
        u0 = self.prob.get("U")
        mask0 = self.prob.fail_mask()
        ni, nt = self.n_int, 3 * (len(self.x) - self.n_int) // 6
        start = u0.copy()
        start[1, ni:ni + nt] += DU
        start[1, ni + nt:] -= DU
        fresh = uses_network(self.k, mask0, self.mask_prev)
        guess = u0 + (self.base if fresh else self.du_prev)
        used_net = False
        if use_net and net is not None and fresh:
            mir, bmir = self.mirror_maps()
            if np.array_equal(mask0, mask0[bmir]):
                used_net = True
                x, ef = self.features(mask0)
                net.set_features(x, ef)
                y_net = net.forward().astype(np.float64).T
                y_odd = 0.5 * (y_net - y_net[:, mir])
                guess[:, :ni] += (NET_SCALE * DU) * y_odd[:, :ni]
        t_pred = time.perf_counter() - t
        guess[:, ni:] = start[:, ni:]
        fell_back = False
        if used_net and fallback is not None:
            fallback.prob.set_fail_mask(mask0)
            its, r, fell_back = race(self, guess, fallback, start, tol, max_iters, chunk)
            if fell_back:
                self.prob.set("U", fallback.prob.get("U"))
                self.prob.set_fail_mask(fallback.prob.fail_mask())
        elif used_net and probe > 0:
            its, r, fell_back = probe_then_solve(self.prob, guess, start, mask0, probe, tol, max_iters)
        else:
            self.prob.set("U", guess)
            its, r, ok = self.prob.adr_solve(0, max_iters, tol, nonfinite_ok=True)
            if not ok:
                fell_back = True
                self.prob.set_fail_mask(mask0)
                self.prob.set("U", start)
                its2, r, _ = self.prob.adr_solve(0, max_iters, tol)
                its += its2
        self.prob.sync()
        sec = time.perf_counter() - t
        self.du_prev = self.prob.get("U") - u0
        self.mask_prev = mask0
        self.k += 1
        return its, r, sec, fell_back, t_pred, used_net


def probe_then_solve(prob, guess, start, mask0, probe, tol, max_iters):
    """The fallback rule of the warm-start configuration: `probe` iterations
    of plain ADR from the plain start, then `probe` iterations from the guess
    (fail mask reset to mask0 before each run); when the guess's residual is
    the larger (or not finite), plain ADR runs again from the plain start to
    tol, otherwise the run from the guess continues to tol.  A probe that
    reaches tol ends the increment.  Every iteration of every run is
    counted.  Returns (iterations, residual, fell_back)."""
    prob.set("U", start)
    ip, rp, okp = prob.adr_solve(0, probe, tol, nonfinite_ok=True)
    if okp:
        return ip, rp, True
    prob.set_fail_mask(mask0)
    prob.set("U", guess)
    iw, rw, okw = prob.adr_solve(0, probe, tol, nonfinite_ok=True)
    if okw:
        return ip + iw, rw, False
    if rw <= rp:
        i2, r2, ok2 = prob.adr_solve(iw, max_iters - iw, tol, nonfinite_ok=True)
        if ok2:
            return ip + iw + i2, r2, False
        iw += i2
    prob.set_fail_mask(mask0)
    prob.set("U", start)
    i3, r3, _ = prob.adr_solve(0, max_iters, tol)
    return ip + iw + i3, r3, True


def race(a, guess_a, b, guess_b, tol, max_iters, chunk=256):
    """ADR from guess_a on plate a and from guess_b on plate b at once (two
    handles, two streams, two host threads; the C-ABI calls release the
    interpreter lock).  Each runs chunks of `chunk` iterations and stops at
    the first chunk boundary after either has reached tol.  Returns
    (iterations, residual, b_won) of the run that finished first (a's when
    neither reached tol)."""
    import threading
    done = threading.Event()
    out = {}

    def run(name, P, g):
        P.prob.set("U", g)
        its, r, ok = 0, float("inf"), False
        while its < max_iters and not done.is_set():
            m = min(chunk, max_iters - its)
            it, r, ok = P.prob.adr_solve(its, m, tol)
            its += it
            if ok or it < m:
                done.set()
                break
        out[name] = (its, r, ok)

    tb = threading.Thread(target=run, args=("b", b, guess_b))
    tb.start()
    run("a", a, guess_a)
    tb.join()
    ia, ra, oka = out["a"]
    ib, rb, okb = out["b"]
    if okb and (not oka or ib < ia):
        return ib, rb, True
    return ia, ra, False


def eval_specs(count, n_min=200, n_max=400, seed_base=0x5EED1004):
    return PB.plate_family(count, seed_base=seed_base, n_min=n_min, n_max=n_max, dx=DX)


def evaluate(net, specs, n_inc, tol=1e-6, probe=50, log=None, race_fallback=False, net_probe=50):
    """Plain ADR, the warm start without the network (du_prev, uniform stretch)
    and with it, on the same plates and increments (three handles per plate,
    each on its own load path; with race_fallback a fourth runs the network
    path's plain-ADR fallback at the same time; with net_probe > 0 the
    network's guess is checked by probe_then_solve; see Plate.step_warm).  The
    network run's seconds include its prediction and its fallback.
    Returns per-increment rows (spec, k, its_plain, its_extrap, its_net, t_plain,
    t_extrap, t_net, fallback_extrap, fallback_net, used_net)."""
    rows = []
    for s, (n, _, clen, ang) in enumerate(specs):
        plain, extrap, warm = Plate(n, clen, ang), Plate(n, clen, ang), Plate(n, clen, ang)
        shadow = Plate(n, clen, ang) if race_fallback else None
        net.set_graph(warm.rp, warm.col)
        try:
            for k in range(n_inc):
                ip, rp_, rprobe, tp = plain.step_plain(tol, probe=probe)
                ie, re_, te, fe, _, _ = extrap.step_warm(None, tol, use_net=False)
                iw, rw, tw, fw, tpred, used = warm.step_warm(net, tol, use_net=True, fallback=shadow,
                                                             probe=net_probe)
                rows.append((s, k, ip, ie, iw, tp, te, tw, fe, fw, used))
                if log:
                    log(f"[eval] plate {s} n={n} inc {k}: iterations plain {ip} no-net {ie} net {iw}  "
                        f"time {tp*1e3:.1f}/{te*1e3:.1f}/{tw*1e3:.1f} ms (predict {tpred*1e3:.1f} ms)"
                        f"{' [network]' if used else ''}{' fallback-net' if fw else ''}"
                        f"{' fallback-extrap' if fe else ''}")
        finally:
            plain.close()
            extrap.close()
            warm.close()
            if shadow is not None:
                shadow.close()
    return rows


def main(argv=None):
    """Evaluate the warm start on unseen plates:
    python -m pd_gpu1_amd.warmstart --plates 128 --increments 6
    prints one line per increment and a JSON summary (totals of iterations and
    seconds for plain ADR, the warm start without and with the network)."""
    import argparse
    import json
    import os

    ap = argparse.ArgumentParser(description="warm-start evaluation on unseen plates")
    ap.add_argument("--plates", type=int, default=128)
    ap.add_argument("--first", type=int, default=0, help="index of the first plate of the seeded family")
    ap.add_argument("--increments", type=int, default=6)
    ap.add_argument("--race", action="store_true", help="run the plain-ADR fallback beside the network's start")
    ap.add_argument("--probe", type=int, default=50,
                    help="check the network's start by the residual after this many iterations against plain ADR's "
                         "(probe_then_solve; 0: restart fallback only)")
    ap.add_argument("--weights", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                                      "mpnn_warmstart.npy"))
    ap.add_argument("--eager-iterations", type=int, default=0,
                    help="only run this many plain ADR iterations of the first plate's first increment, one "
                         "pd_adr_iterate call each (no captured graph; for a per-kernel trace), and print the time")
    a = ap.parse_args(argv)
    specs = eval_specs(a.first + a.plates)[a.first:]
    if a.eager_iterations > 0:
        n, _, clen, ang = specs[0]
        plate = Plate(n, clen, ang)
        try:
            plate.prob.strip_displace(DU, -DU)
            plate.prob.adr_iterate(0)
            plate.prob.sync()
            t0 = time.perf_counter()
            for ct in range(1, a.eager_iterations + 1):
                plate.prob.adr_iterate(ct, sync=False)
            c, r = plate.prob.adr_iterate(a.eager_iterations + 1)
            sec = time.perf_counter() - t0
        finally:
            plate.close()
        print(json.dumps({"plate": a.first, "side": n, "particles": len(plate.x), "bonds": len(plate.col),
                          "eager_iterations": a.eager_iterations + 1, "us_per_iteration": sec * 1e6 /
                          (a.eager_iterations + 1), "residual": r}), flush=True)
        return 0
    with Mpnn() as net:
        net.set_params(np.load(a.weights))
        rows = np.array(evaluate(net, specs, a.increments, log=lambda m: print(m, flush=True), race_fallback=a.race,
                                 net_probe=a.probe), dtype=float)
    first, later = rows[:, 1] == 0, rows[:, 1] >= 1
    tot = lambda sel, c: float(rows[sel, c].sum())
    every = rows[:, 1] >= 0
    out = {"plates": a.plates, "first": a.first, "race": a.race, "probe": a.probe, "net_scale": NET_SCALE, "sides": [int(s[0]) for s in specs], "increments": a.increments,
           "iterations": {"plain": tot(every, 2), "warm_no_network": tot(every, 3), "warm_network": tot(every, 4)},
           "seconds": {"plain": tot(every, 5), "warm_no_network": tot(every, 6), "warm_network": tot(every, 7)},
           "first_increment_iterations": {"plain": tot(first, 2), "uniform_stretch": tot(first, 3),
                                          "network": tot(first, 4)},
           "later_increment_iterations": {"plain": tot(later, 2), "warm": tot(later, 3)},
           "network_used": int(rows[:, 10].sum()), "network_fell_back": int(rows[first, 9].sum())}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    main()
