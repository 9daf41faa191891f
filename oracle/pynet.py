"""CPU oracle (numpy, FP64) of the learned preconditioner network.

TEST INFRASTRUCTURE ONLY (tests/ and the checker legs); the product runs
the network with its own BF16 MFMA kernels (hydro_amd/csrc/net.hip).

The network is a 4-level U-Net on the cell grid (no reference counterpart),
channels c = (16, 32, 64, 64):
  encoder  E0: 1 -> c0 at level 0; for l = 1..L-1 a 2x2 average pool then
           E_l: c_{l-1} -> c_l; every encoder conv is followed by ReLU;
  decoder  D_{L-1}: c_{L-1} -> c_{L-2} on the top encoder output; for
           l = L-2..1 the nearest 2x upsample of the coarser decoder output
           plus the encoder output of level l feeds D_l: c_l -> c_{l-1}
           (+ ReLU); at level 0 the same sum feeds D0: c0 -> 1 (no ReLU).
All convolutions are 'same' 3x3 with zero padding and no bias, so the
network is positively homogeneous.  Weights W[tap][ci][co], tap = 3*ky+kx,
offsets (kx-1, ky-1); activations are (H, W, C) arrays.
"""
import numpy as np


def channels(levels=4, c0=16, cmax=64):
    ch = [min(c0 * 2 ** l, cmax) for l in range(levels)]
    ch[-1] = ch[-2]
    return ch


def layer_shapes(levels=4, c0=16, cmax=64):
    """(name, cin, cout, level) in the order weights are packed."""
    ch = channels(levels, c0, cmax)
    out = [("E0", 1, ch[0], 0)]
    for l in range(1, levels):
        out.append((f"E{l}", ch[l - 1], ch[l], l))
    for l in range(levels - 1, 0, -1):
        out.append((f"D{l}", ch[l], ch[l - 1], l))
    out.append(("D0", ch[0], 1, 0))
    return out


def num_weights(levels=4, c0=16, cmax=64):
    return sum(9 * ci * co for _, ci, co, _ in layer_shapes(levels, c0, cmax))


def unpack(w, levels=4, c0=16, cmax=64):
    out, off = {}, 0
    for name, ci, co, _ in layer_shapes(levels, c0, cmax):
        n = 9 * ci * co
        out[name] = w[off:off + n].reshape(9, ci, co)
        off += n
    return out


def conv(x, W):
    h, w_, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    out = np.zeros((h, w_, W.shape[2]))
    for t in range(9):
        ky, kx = divmod(t, 3)
        out += xp[ky:ky + h, kx:kx + w_, :] @ W[t]
    return out


def conv_bwd(x, W, g):
    """Gradients of sum(g * conv(x, W)) w.r.t. x and W."""
    h, w_, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    gxp = np.zeros_like(xp)
    gW = np.zeros_like(W)
    for t in range(9):
        ky, kx = divmod(t, 3)
        xs = xp[ky:ky + h, kx:kx + w_, :]
        gW[t] = xs.reshape(-1, xs.shape[2]).T @ g.reshape(-1, g.shape[2])
        gxp[ky:ky + h, kx:kx + w_, :] += g @ W[t].T
    return gxp[1:-1, 1:-1, :], gW


def pool(x):
    return 0.25 * (x[0::2, 0::2] + x[0::2, 1::2] + x[1::2, 0::2] + x[1::2, 1::2])


def pool_bwd(g):
    return 0.25 * np.repeat(np.repeat(g, 2, 0), 2, 1)


def up(x):
    return np.repeat(np.repeat(x, 2, 0), 2, 1)


def up_bwd(g):
    return g[0::2, 0::2] + g[0::2, 1::2] + g[1::2, 0::2] + g[1::2, 1::2]


def bf16(a):
    """Rounds to the nearest bfloat16 (ties to even), via float32; returns
    float64.  The GPU kernels store activations, their gradients and the
    16..64-channel weights in this format."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(a):
    return a


def tie(P, levels=4):
    """Tied decoder: D_l[t] = E_l[8 - t]^T (taps flipped, channels swapped)."""
    Q = dict(P)
    for l in range(levels):
        Q[f"D{l}"] = np.transpose(P[f"E{l}"][::-1], (0, 2, 1))
    return Q


def forward(w, x, levels=4, c0=16, cmax=64, keep=False, rnd=None, linear=False, tied=False):
    """x: (H, W) input; returns (H, W) output (and the activations).

    rnd: rounding applied where the GPU stores a tensor (None: exact FP64;
    `bf16`: the GPU's number format -- BF16 activations and inner-layer
    weights, FP32 first/last-layer weights, input and output).
    linear: no activations.  tied (implies linear): the decoder layers are
    the transposes of the encoder layers, y = sum_l 4^l F_l^T F_l x."""
    R = rnd or _ident
    P = unpack(w, levels, c0, cmax)
    names = [n for n, _, _, _ in layer_shapes(levels, c0, cmax)]
    P = {k: (v if k in (names[0], names[-1]) else R(v)) for k, v in P.items()}
    if tied:
        P = tie(P, levels)
        linear = True
    act = (lambda a: a) if linear else (lambda a: np.maximum(a, 0.0))
    acts = {}
    h = x[:, :, None]
    enc = []
    for l in range(levels):
        if l > 0:
            h = R(pool(h))
        acts[f"in_E{l}"] = h
        h = R(act(conv(h, P[f"E{l}"])))
        enc.append(h)
    d = enc[-1]
    for l in range(levels - 1, 0, -1):
        dl_in = enc[l] if l == levels - 1 else R(up(d) + enc[l])
        acts[f"in_D{l}"] = dl_in
        d = R(act(conv(dl_in, P[f"D{l}"])))
        acts[f"out_D{l}"] = d
    d0_in = R(up(d) + enc[0])
    acts["in_D0"] = d0_in
    y = conv(d0_in, P["D0"])[:, :, 0]
    acts["enc"] = enc
    acts["P"] = P
    return (y, acts) if keep else y


def backward(w, x, gy, levels=4, c0=16, cmax=64, rnd=None, linear=False, tied=False):
    """Gradient of sum(gy * forward(x)) w.r.t. the packed weights and x
    (rnd, linear, tied as in forward; gradients of activations are rounded
    where the GPU stores them; with tied weights the decoder slots get zero
    and their gradient is added to the encoder layer they mirror)."""
    R = rnd or _ident
    linear = linear or tied
    _, A = forward(w, x, levels, c0, cmax, keep=True, rnd=rnd, linear=linear, tied=tied)
    P = A["P"]
    enc = A["enc"]
    G = {}
    g_in, G["D0"] = conv_bwd(A["in_D0"], P["D0"], gy[:, :, None])
    g_in = R(g_in)
    g_enc = [None] * levels
    g_enc[0] = g_in                 # in_D0 = up(out_D1) + enc0
    g_d = R(up_bwd(g_in))           # gradient w.r.t. the output of D1
    for l in range(1, levels):
        g_pre = g_d if linear else g_d * (A[f"out_D{l}"] > 0)
        g_in, G[f"D{l}"] = conv_bwd(A[f"in_D{l}"], P[f"D{l}"], g_pre)
        g_in = R(g_in)
        g_enc[l] = g_in
        if l < levels - 1:
            g_d = R(up_bwd(g_in))
    # encoder, from the top down; the pooled path adds into the finer level
    g_x = None
    for l in range(levels - 1, -1, -1):
        g_pre = g_enc[l] if linear else g_enc[l] * (enc[l] > 0)
        g_x, G[f"E{l}"] = conv_bwd(A[f"in_E{l}"], P[f"E{l}"], g_pre)
        if l > 0:
            g_enc[l - 1] = R(g_enc[l - 1] + pool_bwd(R(g_x)))
    if tied:
        for l in range(levels):
            G[f"E{l}"] = G[f"E{l}"] + np.transpose(G[f"D{l}"], (0, 2, 1))[::-1]
            G[f"D{l}"] = np.zeros_like(G[f"D{l}"])
    gw = np.concatenate([G[name].reshape(-1) for name, _, _, _ in layer_shapes(levels, c0, cmax)])
    return gw, g_x[:, :, 0]


def init_weights(levels=4, c0=16, cmax=64, seed=0, scale=1.0):
    """He-style initialisation; the last layer starts at zero so the
    untrained network adds nothing."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, ci, co, _ in layer_shapes(levels, c0, cmax):
        std = scale * np.sqrt(2.0 / (9 * ci))
        wl = rng.normal(0.0, std, (9, ci, co))
        if name == "D0":
            wl[:] = 0.0
        parts.append(wl.reshape(-1))
    return np.concatenate(parts)
