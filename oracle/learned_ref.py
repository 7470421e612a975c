"""CPU restatement of the learned warm start's network (include/dmtz_learned.h) -- TEST
INFRASTRUCTURE ONLY: it checks the HIP forward and backward kernels and is never used by the
product path.

float64 numpy. `exact=False` inserts the device's bf16 roundings (weights, hidden activations,
output gradient, propagated gradients), so the device results differ from it only by fp32
accumulation order; `exact=True` is the plain float64 network, whose gradient is pinned against
central finite differences in the tests. The reference has no learned part, so there is nothing
of it to follow here.
"""
import numpy as np

TAPS = [(t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) for t in range(27)]  # (dx, dy, dz)
CH = 16
LP = 27 * CH * CH + CH
DIRS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (-1, 1, 0), (1, -1, 0),
        (0, 1, 1), (0, -1, -1), (-1, 0, 1), (1, 0, -1), (-1, 1, 1), (1, -1, -1)]


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def bf16_bits(h):
    """float64 values of bf16 bit patterns (uint16 or int16 array)."""
    u = np.asarray(h).astype(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out = []
    for l in range(4):
        w = p[l * LP:l * LP + 27 * CH * CH].reshape(27, CH, CH)
        b = p[l * LP + 27 * CH * CH:(l + 1) * LP]
        out.append((w, b))
    return out


def _shift(x, dx, dy, dz, sign):
    """x[v + sign * off] with zeros outside the grid; x is [z, y, x, c]."""
    nz, ny, nx = x.shape[:3]
    p = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    sx, sy, sz = sign * dx, sign * dy, sign * dz
    return p[1 + sz:1 + sz + nz, 1 + sy:1 + sy + ny, 1 + sx:1 + sx + nx]


def conv(x, w, sign=1, transpose=False):
    """sum over taps of x[v + sign * off] @ W[tap] (or W[tap]^T)."""
    y = np.zeros(x.shape[:3] + (CH,))
    for t, (dx, dy, dz) in enumerate(TAPS):
        m = w[t].T if transpose else w[t]
        y += _shift(x, dx, dy, dz, sign) @ m
    return y


def wgrad(x, dy):
    dw = np.zeros((27, CH, CH))
    d2 = dy.reshape(-1, CH)
    for t, (dx, dyy, dz) in enumerate(TAPS):
        dw[t] = _shift(x, dx, dyy, dz, 1).reshape(-1, CH).T @ d2
    return dw, d2.sum(axis=0)


def forward(feat, params, exact=False):
    """feat [z, y, x, 16] float64 -> (activations, output [z, y, x, 2])."""
    r = (lambda a: a) if exact else bf16
    layers = unpack(params)
    acts = [feat]
    x = feat
    for l in range(3):
        w, b = layers[l]
        x = r(np.maximum(conv(x, r(w)) + b, 0.0))
        acts.append(x)
    w, b = layers[3]
    z = conv(x, r(w)) + b
    return acts, z[..., :2]


def loss_grad(z, lab, pos_weight, lam):
    """Mean loss terms and d(loss)/d(output) [z, y, x, 2]."""
    n = z[..., 0].size
    zz, m = z[..., 0], z[..., 1]
    y, t = lab[..., 0], lab[..., 1]
    sp_pos = np.logaddexp(0.0, -zz)  # -log sigmoid(z)
    sp_neg = np.logaddexp(0.0, zz)   # -log(1 - sigmoid(z))
    bce = (pos_weight * y * sp_pos + (1 - y) * sp_neg).sum() / n
    size = (y * (m - t) ** 2).sum() / n
    p = 1.0 / (1.0 + np.exp(-zz))
    dz = (pos_weight * y * (p - 1) + (1 - y) * p) / n
    dm = 2 * lam * y * (m - t) / n
    return (bce, size), np.stack([dz, dm], axis=-1)


def gradient(feat, lab, params, pos_weight=1.0, lam=1.0, exact=False):
    """Parameter gradient (flat, like the device's) and the mean loss terms."""
    r = (lambda a: a) if exact else bf16
    layers = unpack(params)
    acts, z = forward(feat, params, exact)
    (bce, size), dout = loss_grad(z, lab, pos_weight, lam)
    g = np.zeros(CH)
    d = np.zeros(z.shape[:3] + (CH,))
    d[..., :2] = r(dout)
    grads = [None] * 4
    for l in range(3, -1, -1):
        w, _ = layers[l]
        dw, db = wgrad(acts[l], d)
        grads[l] = (dw, db)
        if l > 0:
            d = r(conv(d, r(w), sign=-1, transpose=True) * (acts[l] > 0))
    flat = np.concatenate([np.concatenate([dw.ravel(), db]) for dw, db in grads])
    return flat, (bce, size)


def mask():
    """1 for the trainable entries (layer 0 inputs 0..7, layer 3 outputs 0..1), else 0."""
    m = np.zeros(4 * LP)
    for l in range(4):
        cin = 8 if l == 0 else CH
        cout = 2 if l == 3 else CH
        w = np.zeros((27, CH, CH))
        w[:, :cin, :cout] = 1
        b = np.zeros(CH)
        b[:cout] = 1
        m[l * LP:(l + 1) * LP] = np.concatenate([w.ravel(), b])
    return m


def features(g, eps, rng):
    """Input channels 0, 5, 6, 7 (float32 arithmetic in the device's order, then bf16); channels
    1..4 come from the lower-star classifier and are checked against dmtz_classify."""
    g = np.asarray(g, dtype=np.float32)
    nz, ny, nx = g.shape
    out = np.zeros(g.shape + (CH,))
    inv = np.float32(1.0 / eps)
    s = np.zeros(g.shape, np.float32)
    mn = g.copy()
    mx = g.copy()
    cnt = np.zeros(g.shape, np.int32)
    pad = np.pad(g, 1)
    valid = np.pad(np.ones(g.shape, bool), 1)
    for dx, dy, dz in DIRS:
        w = pad[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
        ok = valid[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
        s = np.where(ok, (s + w).astype(np.float32), s)
        mn = np.where(ok & (w < mn), w, mn)
        mx = np.where(ok & (w > mx), w, mx)
        cnt += ok
    mean = np.where(cnt > 0, (s / cnt.astype(np.float32)).astype(np.float32), g)
    out[..., 0] = bf16(np.clip(((g - mean).astype(np.float32) * inv).astype(np.float32), -8, 8))
    out[..., 5] = bf16(np.float32(np.log10(eps / rng) + 3.0))
    out[..., 6] = bf16(np.clip(((g - mn).astype(np.float32) * inv).astype(np.float32), -8, 8))
    out[..., 7] = bf16(np.clip(((mx - g).astype(np.float32) * inv).astype(np.float32), -8, 8))
    return out
