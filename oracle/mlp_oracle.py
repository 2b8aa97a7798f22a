"""CPU restatement (numpy) of the learned warm start's network -- TEST INFRASTRUCTURE ONLY.

Only tests/ use this module, as the checker of the GPU kernels; the product never imports it.
The warm start has no reference counterpart (PhyBS has no learned component): this is the
specification the GPU kernels implement, pinned in tests/test_learn_oracle.py by FP64 central
differences of the loss (the FP64 variant, rounding off) and by the FP32/BF16 variant agreeing
with the FP64 one to BF16 precision.

Network: per vertex, 18 features -> 64 -> 64 -> 64 -> 3, SiLU on the hidden layers.
  features (per free node, step start): u = x - X, h v, s = h v + h^2 g, and the 1-ring means of
  (u, h v, s) minus the node's own values; 6 groups of 3.
  normalised input  xin[k] = bf16(f[k] * inv_s[k // 3])       inv_s = float32(1) / float32(scale)
  z_l = b_l + sum_k bf16(W_l[k, o]) * h_{l-1}[k]  (FP32 accumulation)
  h_l = bf16(silu(z_l)) for l = 1..3; out = z_4
  prediction (displacement from the inertial target y = x + h v + h^2 g) = out * scale[6]
Loss over a batch of B samples: mean over (sample, component) of (out - t * inv_s[6])^2.
Backward in FP32, bf16 rounding of activations passed straight through.
Parameters are packed: W1 (18 x 64, row-major in x out), b1, W2 (64 x 64), b2, W3, b3,
W4 (64 x 3), b4: 9,731 floats.
"""
import numpy as np

SIZES = [(18, 64), (64, 64), (64, 64), (64, 3)]
NPARAM = sum(i * o + o for i, o in SIZES)  # 9731
NSCALE = 7


def offsets():
    out, off = [], 0
    for i, o in SIZES:
        out.append((off, off + i * o, off + i * o + o))  # W start, b start, end
        off += i * o + o
    return out


def unpack(params, dtype=np.float32):
    p = np.asarray(params, dtype)
    Ws, bs = [], []
    for (i, o), (w0, b0, e) in zip(SIZES, offsets()):
        Ws.append(p[w0:b0].reshape(i, o))
        bs.append(p[b0:e])
    return Ws, bs


def bf16(a):
    """Round float32 to the nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def init_params(seed):
    """Glorot-uniform weights, zero biases (test inputs; the GPU takes them through
    pbs_learn_set_model)."""
    rng = np.random.default_rng(seed)
    out = []
    for i, o in SIZES:
        lim = np.sqrt(6.0 / (i + o))
        out.append(rng.uniform(-lim, lim, (i, o)).astype(np.float32).ravel())
        out.append(np.zeros(o, np.float32))
    return np.concatenate(out)


def silu(z):
    return z / (np.float32(1) + np.exp(-z))


def dsilu(z):
    s = np.float32(1) / (np.float32(1) + np.exp(-z))
    return s * (np.float32(1) + z * (np.float32(1) - s))


def inv_scales(scales, dtype=np.float32):
    return dtype(1) / np.asarray(scales, dtype)


def forward(params, scales, feats, rounding=True):
    """Returns (out [B,3] normalised, cache).  rounding=False: FP64, no bf16 (the pinned variant)."""
    dt = np.float32 if rounding else np.float64
    r = bf16 if rounding else (lambda a: a)
    Ws, bs = unpack(params, dt)
    inv = inv_scales(scales, dt)
    f = np.asarray(feats, dt)
    xin = r((f * np.repeat(inv[:6], 3)[None, :]).astype(dt))
    hs, zs = [xin], []
    h = xin
    for li, (W, b) in enumerate(zip(Ws, bs)):
        z = (b.astype(dt)[None, :] + h @ r(W.astype(dt))).astype(dt)
        zs.append(z)
        if li < 3:
            h = r(silu(z).astype(dt))
            hs.append(h)
    return zs[-1], (hs, zs, inv)


def loss_grad(params, scales, feats, targets, rounding=True):
    """Loss and gradient (packed like params)."""
    dt = np.float32 if rounding else np.float64
    r = bf16 if rounding else (lambda a: a)
    out, (hs, zs, inv) = forward(params, scales, feats, rounding)
    B = out.shape[0]
    t = np.asarray(targets, dt) * inv[6]
    diff = out - t
    loss = float(np.mean(diff.astype(np.float64) ** 2))
    d = (dt(2) / dt(3 * B)) * diff
    Ws, bs = unpack(params, dt)
    grads_W, grads_b = [None] * 4, [None] * 4
    for li in range(3, -1, -1):
        grads_W[li] = (hs[li].T @ d).astype(dt)
        grads_b[li] = d.sum(axis=0).astype(dt)
        if li > 0:
            dh = d @ r(Ws[li].astype(dt)).T
            d = (dh * dsilu(zs[li - 1])).astype(dt)
    g = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in zip(grads_W, grads_b)])
    return loss, g.astype(np.float64 if not rounding else np.float32)


def features(x, v, rest, h, grav, adj_off, adj, mask):
    """Per-node features (float64 in, float32 out) in node order; rows of fixed nodes are zero.
    adj_off/adj: node adjacency (CSR, neighbours without the node itself)."""
    x = np.asarray(x, np.float64)
    u = x - np.asarray(rest, np.float64)
    hv = h * np.asarray(v, np.float64)
    s = hv + h * h * np.asarray(grav, np.float64)[None, :]
    n = len(x)
    cnt = np.diff(adj_off)
    rows = np.repeat(np.arange(n), cnt)
    out = np.zeros((n, 18))
    for gi, a in enumerate((u, hv, s)):
        out[:, 3 * gi:3 * gi + 3] = a
        m = np.zeros((n, 3))
        np.add.at(m, rows, a[adj])
        m /= np.maximum(cnt, 1)[:, None]
        out[:, 9 + 3 * gi:12 + 3 * gi] = m - a
    out[np.asarray(mask) != 0] = 0
    return out.astype(np.float32)
