"""CPU restatement of the learned moment denoiser (swap_amd/csrc/k_denoise.hip).

TEST INFRASTRUCTURE ONLY: the checker of the device's forward and backward
kernels.  The net has no counterpart in the reference; this is its
mathematical definition: 4 -> 32 -> 32 -> 32 -> 4 channels, 3x3x3 kernels
with periodic padding (tap = dx+1 + 3(dy+1) + 9(dz+1)), leaky ReLU (slope
LEAK = 1/16 below zero) after layers 1-3, residual output, mean squared
error.  With bf16=True every tensor the
device stores in BF16 (activations, weights, gradients passed between
layers) is rounded the same way (round to nearest even) and every sum is
taken in float64; with bf16=False it is the exact float64 network.

Parameters are flat float32 in layer order: W[l][tap][c_in][c_out], b[l][c_out].
Grids are [V][C] with the voxel index v = x + nx (y + ny z).
"""
import numpy as np

CH = (4, 32, 32, 32, 4)
NL = 4
LEAK = 0.0625


def n_params():
    return sum(27 * CH[l] * CH[l + 1] + CH[l + 1] for l in range(NL))


def unpack(params):
    out, o = [], 0
    for l in range(NL):
        nw = 27 * CH[l] * CH[l + 1]
        W = np.asarray(params[o:o + nw], np.float64).reshape(27, CH[l], CH[l + 1])
        o += nw
        b = np.asarray(params[o:o + CH[l + 1]], np.float64)
        o += CH[l + 1]
        out.append((W, b))
    return out


def bf16(x):
    """float -> nearest BF16 (ties to even), returned as float64"""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def neighbours(shape):
    """nbr[v, t]: voxel index of v shifted by tap t, periodic"""
    nx, ny, nz = shape
    v = np.arange(nx * ny * nz)
    x, y, z = v % nx, (v // nx) % ny, v // (nx * ny)
    t = np.arange(27)
    dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
    a = (x[:, None] + dx[None]) % nx
    b = (y[:, None] + dy[None]) % ny
    c = (z[:, None] + dz[None]) % nz
    return a + nx * (b + ny * c)


def conv(X, W, nbr):
    """Y[v, co] = sum_t sum_ci W[t, ci, co] X[nbr[v, t], ci]"""
    return np.einsum("vtc,tcd->vd", X[nbr], W)


def conv_adj(G, W, nbr):
    """adjoint of conv: D[v, ci] = sum_t sum_co W[t, ci, co] G[nbr[v, 26 - t], co]"""
    return np.einsum("vtd,tcd->vc", G[nbr[:, ::-1]], W)


def wgrad(X, G, nbr):
    """dW[t, ci, co] = sum_v X[nbr[v, t], ci] G[v, co]"""
    return np.einsum("vtc,vd->tcd", X[nbr], G)


def forward_backward(params, x, target, shape, bf=True):
    """(loss, out, grads flat) of one forward + backward; x, target [V][4]"""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params)
    nbr = neighbours(shape)
    x = np.asarray(x, np.float64)
    acts = [r(x)]
    Ws = [r(W) for W, _ in layers]
    for l in range(NL - 1):
        z = conv(acts[-1], Ws[l], nbr) + layers[l][1]
        acts.append(r(np.where(z > 0, z, LEAK * z)))
    out = conv(acts[-1], Ws[NL - 1], nbr) + layers[NL - 1][1] + x
    n = out.size
    d = out - target
    loss = float(np.sum(d * d) / n)
    g = r(2.0 * d / n)
    grads = [None] * NL
    for l in range(NL - 1, -1, -1):
        grads[l] = (wgrad(acts[l], g, nbr), g.sum(0))
        if l > 0:
            a = conv_adj(g, Ws[l], nbr)
            g = r(np.where(acts[l] > 0, a, LEAK * a))
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, out, flat
