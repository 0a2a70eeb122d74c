"""FP64 numpy reference of the message-passing warm-start network (test
infrastructure only; the GPU network lives in pd_gpu1_amd/csrc/k_mpnn.hip).

Network (no reference counterpart; the warm-start network of pd_gpu1_amd):
    h0 = silu(x W0 + b0)                                   x: [N, 16]
    for l in 1..3:
        P   = h W1_l                                       [N, 64]
        pre = P[col] + e Q_l + c_l                         per directed edge
        a_i = mean_{edges of i} silu(pre)                  [N, 64]
        h   = h + silu([h, a] W2_l + b2_l)
    y = h Wd + bd                                          [N, 2]
loss = sum_i w_i |y_i - t_i|^2 / sum_i w_i

Parameters are one flat float vector in the order W0, b0, then per layer
W1, Q, c, W2, b2, then Wd, bd; matrices are [in][out] row-major.
"""
import numpy as np

F_IN, E_IN, HID, N_LAYERS, OUT = 16, 3, 64, 3, 2


def shapes():
    s = [("W0", (F_IN, HID)), ("b0", (HID,))]
    for l in range(N_LAYERS):
        s += [(f"W1_{l}", (HID, HID)), (f"Q_{l}", (E_IN, HID)), (f"c_{l}", (HID,)),
              (f"W2_{l}", (2 * HID, HID)), (f"b2_{l}", (HID,))]
    s += [("Wd", (HID, OUT)), ("bd", (OUT,))]
    return s


def n_params():
    return sum(int(np.prod(sh)) for _, sh in shapes())


def unpack(theta):
    out, o = {}, 0
    for name, sh in shapes():
        n = int(np.prod(sh))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(sh)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in shapes()])


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sig(z)


def dsilu(z):
    s = sig(z)
    return s * (1.0 + z * (1.0 - s))


def forward(theta, row_ptr, col, efeat, x, keep=False):
    p = unpack(theta)
    N = len(row_ptr) - 1
    deg = np.diff(row_ptr).astype(np.float64)
    rows = np.repeat(np.arange(N), np.diff(row_ptr))
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    u0 = x @ p["W0"] + p["b0"]
    h = silu(u0)
    cache = {"u0": u0, "layers": []}
    for l in range(N_LAYERS):
        P = h @ p[f"W1_{l}"]
        pre = P[col] + efeat @ p[f"Q_{l}"] + p[f"c_{l}"]
        m = silu(pre)
        a = np.zeros((N, HID))
        np.add.at(a, rows, m)
        a *= inv[:, None]
        z = np.concatenate([h, a], axis=1) @ p[f"W2_{l}"] + p[f"b2_{l}"]
        cache["layers"].append((h, P, pre, a, z))
        h = h + silu(z)
    y = h @ p["Wd"] + p["bd"]
    cache["h_last"] = h
    return (y, cache) if keep else y


def loss_and_grad(theta, row_ptr, col, efeat, x, target, weight):
    p = unpack(theta)
    N = len(row_ptr) - 1
    deg = np.diff(row_ptr).astype(np.float64)
    rows = np.repeat(np.arange(N), np.diff(row_ptr))
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    y, cache = forward(theta, row_ptr, col, efeat, x, keep=True)
    wsum = float(np.sum(weight))
    r = y - target
    loss = float(np.sum(weight[:, None] * r * r) / wsum)
    g = {}
    dy = 2.0 * weight[:, None] * r / wsum
    h = cache["h_last"]
    g["Wd"] = h.T @ dy
    g["bd"] = dy.sum(0)
    dh = dy @ p["Wd"].T
    for l in reversed(range(N_LAYERS)):
        h_in, P, pre, a, z = cache["layers"][l]
        dz = dh * dsilu(z)
        W2 = p[f"W2_{l}"]
        g[f"W2_{l}"] = np.concatenate([h_in, a], axis=1).T @ dz
        g[f"b2_{l}"] = dz.sum(0)
        dh_in = dh + dz @ W2[:HID].T
        da = dz @ W2[HID:].T
        dpre = (da * inv[:, None])[rows] * dsilu(pre)
        g[f"Q_{l}"] = efeat.T @ dpre
        g[f"c_{l}"] = dpre.sum(0)
        dP = np.zeros((N, HID))
        np.add.at(dP, col, dpre)
        g[f"W1_{l}"] = h_in.T @ dP
        dh = dh_in + dP @ p[f"W1_{l}"].T
    du0 = dh * dsilu(cache["u0"])
    g["W0"] = x.T @ du0
    g["b0"] = du0.sum(0)
    return loss, pack(g)
