"""FP64 numpy restatement of the learned preconditioner network (test
infrastructure: only tests/ use it, as the checker of the HIP U-Net kernels).

The network is the 3-level 3D U-Net of the learned right preconditioner:
widths 16/32/64, 8 input channels (residual re/im x 3, log-contrast and
the log cell size in background wavenumbers), 6 output channels, 3x3x3 convolutions
with bias and ReLU, 2x2x2 average pooling, nearest-neighbour upsampling and
channel concatenation [upsampled, skip].  There is no counterpart in the
reference (its preconditioner is the inner GMRES of GFGMRES/seq_fgmres.F90).

Conventions shared with giem2g_amd/csrc/g2g_unet.hip:
  * volumes are channels-last, shape (Nz, Ny, Nx, C), x fastest in space;
  * tap t = (dz+1)*9 + (dy+1)*3 + (dx+1), offsets in {-1, 0, 1}, zero padding;
  * weights W[co, t, ci] and bias b[co] in FP32, flattened layer by layer
    (W then b) into one parameter vector;
  * BF16 storage: weights and every activation that feeds a convolution are
    rounded to bfloat16 (round to nearest even) before use; accumulation is
    exact here (FP32 on the GPU).
"""
import numpy as np

# (name, Cin, Cout, taps, relu)
LAYERS = [("e1a", 8, 16, 27, True), ("e1b", 16, 16, 27, True),
          ("e2a", 16, 32, 27, True), ("e2b", 32, 32, 27, True),
          ("e3a", 32, 64, 27, True), ("e3b", 64, 64, 27, True),
          ("d2a", 96, 32, 27, True), ("d2b", 32, 32, 27, True),
          ("d1a", 48, 16, 27, True), ("d1b", 16, 16, 27, True),
          ("out", 16, 6, 1, False)]


def n_params():
    return sum(co * t * ci + co for _, ci, co, t, _ in LAYERS)


def split(params):
    """Parameter vector -> list of (W[co,t,ci], b[co])."""
    out, o = [], 0
    for _, ci, co, t, _ in LAYERS:
        W = params[o:o + co * t * ci].reshape(co, t, ci)
        o += co * t * ci
        b = params[o:o + co]
        o += co
        out.append((W, b))
    return out


def init_params(seed):
    """He-normal weights, zero bias; the output layer is scaled by 0.1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ps = []
    for name, ci, co, t, _ in LAYERS:
        std = np.sqrt(2.0 / (ci * t)) * (0.1 if name == "out" else 1.0)
        ps.append((rng.standard_normal((co, t, ci)) * std).ravel())
        ps.append(np.zeros(co))
    return np.concatenate(ps).astype(np.float32)


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _taps(ntap):
    if ntap == 1:
        return [(0, 0, 0)]
    return [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def _shift(x, dz, dy, dx):
    """s[z,y,x] = x[z+dz, y+dy, x+dx] with zeros outside."""
    Z, Y, X = x.shape[:3]
    s = np.zeros_like(x)
    zs, zd = slice(max(dz, 0), Z + min(dz, 0)), slice(max(-dz, 0), Z + min(-dz, 0))
    ys, yd = slice(max(dy, 0), Y + min(dy, 0)), slice(max(-dy, 0), Y + min(-dy, 0))
    xs, xd = slice(max(dx, 0), X + min(dx, 0)), slice(max(-dx, 0), X + min(-dx, 0))
    s[zd, yd, xd] = x[zs, ys, xs]
    return s


def conv(x, W, b):
    """y[p, co] = b[co] + sum_t sum_ci W[co,t,ci] x[p + off(t), ci]."""
    y = np.zeros(x.shape[:3] + (W.shape[0],)) + b
    for t, (dz, dy, dx) in enumerate(_taps(W.shape[1])):
        y += _shift(x, dz, dy, dx) @ W[:, t, :].T
    return y


def conv_back(x, W, dy, qdx=None):
    """Gradients of conv: (dx, dW, db) for upstream gradient dy; qdx, when
    given, rounds dy before the data gradient (the GPU feeds a bfloat16 copy
    of dy to the MFMA data-gradient convolution)."""
    dx = np.zeros(x.shape)
    dW = np.zeros(W.shape)
    dyq = dy if qdx is None else qdx(dy)
    for t, (dz, dyy, ddx) in enumerate(_taps(W.shape[1])):
        xs = _shift(x, dz, dyy, ddx)
        dW[:, t, :] = np.einsum("zyxo,zyxi->oi", dy, xs)
        dx += _shift(dyq @ W[:, t, :], -dz, -dyy, -ddx)
    return dx, dW, dy.sum(axis=(0, 1, 2))


def pool(x):
    Z, Y, X, C = x.shape
    return x.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).mean(axis=(1, 3, 5))


def pool_back(d):
    return np.repeat(np.repeat(np.repeat(d, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_back(d):
    Z, Y, X, C = d.shape
    return d.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).sum(axis=(1, 3, 5))


def forward(params, x, round_bf16=True):
    """x: (Nz, Ny, Nx, 8) -> (out (Nz, Ny, Nx, 6), cache).  With round_bf16
    the weights and every convolution input are rounded to bfloat16."""
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    P = [(q(W), np.asarray(b, np.float64)) for W, b in split(params)]
    cache = {"P": P}

    def layer(i, inp):
        W, b = P[i]
        xin = q(inp)
        y = conv(xin, W, b)
        if LAYERS[i][4]:
            y = np.maximum(y, 0.0)
        cache[i] = (xin, y)
        return y

    h = layer(0, x)
    e1 = layer(1, h)
    h = layer(2, pool(q(e1)))
    e2 = layer(3, h)
    h = layer(4, pool(q(e2)))
    e3 = layer(5, h)
    h = layer(6, np.concatenate([up(e3), e2], axis=3))
    d2 = layer(7, h)
    h = layer(8, np.concatenate([up(d2), e1], axis=3))
    d1 = layer(9, h)
    out = layer(10, d1)
    return out, cache


def backward(params, cache, dout, round_bf16=False):
    """Gradient of sum(dout * out) w.r.t. the parameter vector.  With
    round_bf16 the data-gradient products use bfloat16-rounded upstream
    gradients, as the GPU does."""
    P = cache["P"]
    grads = [None] * len(LAYERS)
    qdx = bf16 if round_bf16 else None

    def layer_back(i, dy):
        xin, y = cache[i]
        if LAYERS[i][4]:
            dy = dy * (y > 0)
        W, _ = P[i]
        dx, dW, db = conv_back(xin, W, dy, qdx)
        grads[i] = (dW, db)
        return dx

    dd1 = layer_back(10, dout)
    dh = layer_back(9, dd1)
    dcat = layer_back(8, dh)
    c2 = LAYERS[7][2]
    dd2 = up_back(dcat[..., :c2])
    de1 = dcat[..., c2:]
    dh = layer_back(7, dd2)
    dcat = layer_back(6, dh)
    c3 = LAYERS[5][2]
    de3 = up_back(dcat[..., :c3])
    de2 = dcat[..., c3:]
    dh = layer_back(5, de3)
    dp = layer_back(4, dh)
    de2 = de2 + pool_back(dp)
    dh = layer_back(3, de2)
    dp = layer_back(2, dh)
    de1 = de1 + pool_back(dp)
    dh = layer_back(1, de1)
    layer_back(0, dh)
    return np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
