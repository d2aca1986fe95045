"""float64 numpy restatement of the learned sshex8 smoother (test infrastructure).

The smoother has no counterpart in the reference (chosen form 4).  It acts per macro
element on the (L+1)^3 micro-grid, zero padding outside the element:

    g(z) = conv_n( relu( ... relu( conv_1(z) ) ... ) )      3x3x3 convolutions, no biases
    f(z) = g(z) - g(-z)                                      odd, so f(a z) = a f(z) for all a

Input z = D^-1 r restricted to the element, output: the element's correction; one channel for a
scalar problem ([E][n] arrays), or one channel per component for a vector problem (the 3-channel
network of linear elasticity, [E][3][n] arrays).  Weights W_l[c_out][c_in][27] with offset
k = (dz+1)*9 + (dy+1)*3 + (dx+1), node index lidx = z*(L+1)^2 + y*(L+1) + x.  This module gives forward, backward (weights and input) and the
normalised squared-error loss; tests/test_oracle_cnn.py pins it by finite differences, and the
GPU kernels are checked against it.
"""
import numpy as np

OFFSETS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def shift(a, L, dz, dy, dx):
    """b[..., z, y, x] = a[..., z+dz, y+dy, x+dx] with zero padding; a: [..., n], n = (L+1)^3."""
    n1 = L + 1
    g = a.reshape(a.shape[:-1] + (n1, n1, n1))
    out = np.zeros_like(g)
    zs = slice(max(0, -dz), n1 - max(0, dz))
    ys = slice(max(0, -dy), n1 - max(0, dy))
    xs = slice(max(0, -dx), n1 - max(0, dx))
    zt = slice(max(0, dz), n1 - max(0, -dz))
    yt = slice(max(0, dy), n1 - max(0, -dy))
    xt = slice(max(0, dx), n1 - max(0, -dx))
    out[..., zs, ys, xs] = g[..., zt, yt, xt]
    return out.reshape(a.shape)


def conv(x, W, L):
    """x: [E][Cin][n], W: [Cout][Cin][27] -> [E][Cout][n]"""
    out = np.zeros((x.shape[0], W.shape[0], x.shape[2]))
    for k, (dz, dy, dx) in enumerate(OFFSETS):
        xs = shift(x, L, dz, dy, dx)
        out += np.einsum("oc,ecn->eon", W[:, :, k], xs)
    return out


def conv_backward(x, W, dout, L):
    """gradients of conv: (dW [Cout][Cin][27], dx [E][Cin][n])"""
    dW = np.zeros_like(W)
    dx = np.zeros_like(x)
    for k, (dz, dy, dx_) in enumerate(OFFSETS):
        xs = shift(x, L, dz, dy, dx_)
        dW[:, :, k] = np.einsum("eon,ecn->oc", dout, xs)
        # out[n] += W x[n+d]  ->  dx[m] += W^T dout[m-d]
        dx += shift(np.einsum("oc,eon->ecn", W[:, :, k], dout), L, -dz, -dy, -dx_)
    return dW, dx


def g_forward(z, Ws, L):
    """z: [E][n] (one channel) or [E][C][n]; returns the output of the same shape and the cache of
    the layer inputs."""
    a = z[:, None, :] if z.ndim == 2 else z
    cache = []
    for i, W in enumerate(Ws):
        inp = a if i == 0 else np.maximum(a, 0.0)
        cache.append(inp)
        a = conv(inp, W, L)
    return (a[:, 0, :] if z.ndim == 2 else a), cache


def g_backward(dout, Ws, cache, L):
    grads = [None] * len(Ws)
    d = dout[:, None, :] if dout.ndim == 2 else dout
    for i in reversed(range(len(Ws))):
        inp = cache[i]
        dW, dinp = conv_backward(inp, Ws[i], d, L)
        grads[i] = dW
        if i > 0:
            d = dinp * (inp > 0)  # inp = relu(pre); relu' on the pre-activation
        else:
            d = dinp
    return grads, (d[:, 0, :] if dout.ndim == 2 else d)


def forward(z, Ws, L):
    gp, cp = g_forward(z, Ws, L)
    gm, cm = g_forward(-z, Ws, L)
    return gp - gm, (cp, cm)


def loss_and_grads(z, target, Ws, L):
    """loss = sum (f(z) - t)^2 / sum t^2 and its gradients w.r.t. the weights."""
    out, (cp, cm) = forward(z, Ws, L)
    norm = float(np.sum(target ** 2))
    diff = out - target
    loss = float(np.sum(diff ** 2)) / norm
    dout = 2.0 * diff / norm
    gp, _ = g_backward(dout, Ws, cp, L)
    gm, _ = g_backward(-dout, Ws, cm, L)
    return loss, [a + b for a, b in zip(gp, gm)]


def init_weights(hidden=16, layers=4, seed=0, channels=1):
    """He-style initialisation scaled for the 27-point stencils; `channels` input and output
    channels (1, or 3 for the vector network)."""
    rng = np.random.default_rng(seed)
    chans = [channels] + [hidden] * (layers - 1) + [channels]
    Ws = []
    for i in range(layers):
        fan_in = chans[i] * 27
        Ws.append(rng.normal(0, np.sqrt(2.0 / fan_in), (chans[i + 1], chans[i], 27)))
    return Ws


def n_params(hidden=16, layers=4, channels=1):
    chans = [channels] + [hidden] * (layers - 1) + [channels]
    return sum(27 * chans[i] * chans[i + 1] for i in range(layers))
