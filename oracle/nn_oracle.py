"""CPU oracle of the learned preconditioner's network (numpy).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of libteahip.so's
network kernels (exploringsycl_amd/csrc/nn.hip); nothing in the product
imports it.  No counterpart in the reference.

Network: six 3x3 convolutions with zero padding, 32 hidden channels, ReLU
after the first five, 3 input channels, 1 output channel; parameters in the
order W1 (32,3,3,3), b1, W2..W5 (32,32,3,3), b2..b5, W6 (1,32,3,3), b6 with
W[out][in][ky][kx] (tap (ky, kx) reads the input at (y + ky - 1, x + kx - 1)).

`bf16=True` rounds where the GPU stores BF16 (weights, the packed input, the
hidden activations and the backward deltas), round to nearest even;
accumulation is float64 here (FP32 on the GPU), so results differ from the
GPU only by accumulation order and the BF16 roundings that order flips.
`bf16=False` is exact float64 and is what the finite-difference check pins.
"""

import numpy as np

C, CIN, LAYERS = 32, 3, 6
SHAPES = [(C, CIN)] + [(C, C)] * 4 + [(1, C)]


def num_params():
    return sum(o * i * 9 + o for o, i in SHAPES)


def unpack(params):
    """List of (W [o][i][3][3], b [o]) from the flat parameter vector."""
    out, k = [], 0
    for o, i in SHAPES:
        w = params[k:k + o * i * 9].reshape(o, i, 3, 3)
        k += o * i * 9
        b = params[k:k + o]
        k += o
        out.append((w, b))
    return out


def init_params(seed=0):
    """He-normal weights, zero biases, output layer scaled down."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (o, i) in enumerate(SHAPES):
        std = np.sqrt(2.0 / (i * 9)) * (0.1 if li == LAYERS - 1 else 1.0)
        parts.append(rng.standard_normal(o * i * 9) * std)
        parts.append(np.zeros(o))
    return np.concatenate(parts).astype(np.float32)


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def _pad(a):
    return np.pad(a, ((0, 0), (1, 1), (1, 1), (0, 0)))


def conv3x3(a, w):
    """a [B][H][W][i] (float64), w [o][i][3][3] -> [B][H][W][o], zero padding."""
    B, H, W, _ = a.shape
    ap = _pad(a)
    out = np.zeros((B, H, W, w.shape[0]))
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("bhwi,oi->bhwo", ap[:, ky:ky + H, kx:kx + W, :],
                             w[:, :, ky, kx], optimize=True)
    return out


def conv3x3_transpose(d, w):
    """Adjoint of conv3x3 in its input: d [B][H][W][o] -> [B][H][W][i]."""
    B, H, W, _ = d.shape
    dp = _pad(d)
    out = np.zeros((B, H, W, w.shape[1]))
    for ky in range(3):
        for kx in range(3):
            # input pixel p feeds output p - (ky-1, kx-1)
            out += np.einsum("bhwo,oi->bhwi",
                             dp[:, 2 - ky:2 - ky + H, 2 - kx:2 - kx + W, :],
                             w[:, :, ky, kx], optimize=True)
    return out


def conv3x3_wgrad(d, a):
    """dW[o][i][ky][kx] = sum_p d[p][o] a[p + (ky-1, kx-1)][i]."""
    B, H, W, _ = d.shape
    ap = _pad(a)
    g = np.zeros((d.shape[3], a.shape[3], 3, 3))
    for ky in range(3):
        for kx in range(3):
            g[:, :, ky, kx] = np.einsum("bhwo,bhwi->oi", d, ap[:, ky:ky + H, kx:kx + W, :],
                                        optimize=True)
    return g


def forward(params, x, bf16_mode=True):
    """Returns (output [B][H][W], cache)."""
    rnd = bf16 if bf16_mode else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float64 if not bf16_mode else np.float32))
    a = rnd(x).astype(np.float64)
    acts = [a]
    for li in range(LAYERS - 1):
        w, b = layers[li]
        z = conv3x3(a, rnd(w).astype(np.float64)).astype(np.float32 if bf16_mode else np.float64)
        z = z + (np.float32(1) * b if bf16_mode else b)
        a = rnd(np.maximum(z, 0)).astype(np.float64)
        acts.append(a)
    w6, b6 = layers[-1]
    w6r = (bf16(w6) if bf16_mode else w6).astype(np.float64)
    out = conv3x3(a, w6r)[..., 0] + b6[0]
    if bf16_mode:
        out = out.astype(np.float32)
    return out, (acts, layers, rnd)


def loss_grad(params, x, target, bf16_mode=True):
    """Mean squared error and its gradient (flat, master order)."""
    out, (acts, layers, rnd) = forward(params, x, bf16_mode)
    n = out.size
    diff = out.astype(np.float64) - target
    loss = float(np.mean(diff * diff))
    dout = (2.0 * diff / n)
    if bf16_mode:
        dout = dout.astype(np.float32).astype(np.float64)
    grads = [None] * LAYERS
    w6, _ = layers[-1]
    w6r = (bf16(w6) if bf16_mode else w6).astype(np.float64)
    d = dout[..., None]
    grads[-1] = (conv3x3_wgrad(d, acts[-1]), d.sum(axis=(0, 1, 2)))
    delta = conv3x3_transpose(d, w6r) * (acts[-1] > 0)
    delta = rnd(delta).astype(np.float64)
    for li in range(LAYERS - 2, -1, -1):
        w, _ = layers[li]
        grads[li] = (conv3x3_wgrad(delta, acts[li]), delta.sum(axis=(0, 1, 2)))
        if li > 0:
            wr = rnd(w).astype(np.float64)
            delta = conv3x3_transpose(delta, wr) * (acts[li] > 0)
            delta = rnd(delta).astype(np.float64)
    flat = np.concatenate([np.concatenate([g.ravel(), gb.ravel()]) for g, gb in grads])
    return loss, flat


ADAM_BETA1, ADAM_BETA2, ADAM_EPS = 0.9, 0.999, 1e-8


def adam_step(params, grad, m, v, t, lr, beta1=ADAM_BETA1, beta2=ADAM_BETA2,
              eps=ADAM_EPS):
    """One step of Adam (Kingma & Ba 2015, Algorithm 1) in float64.

    t is the 1-based number of this step.  Returns (params, m, v) after it:
        m <- beta1 m + (1 - beta1) g
        v <- beta2 v + (1 - beta2) g^2
        params <- params - lr (m / (1 - beta1^t)) / (sqrt(v / (1 - beta2^t)) + eps)
    """
    g = np.asarray(grad, dtype=np.float64)
    m = beta1 * np.asarray(m, dtype=np.float64) + (1.0 - beta1) * g
    v = beta2 * np.asarray(v, dtype=np.float64) + (1.0 - beta2) * g * g
    m_hat = m / (1.0 - beta1 ** t)
    v_hat = v / (1.0 - beta2 ** t)
    p = np.asarray(params, dtype=np.float64) - lr * m_hat / (np.sqrt(v_hat) + eps)
    return p, m, v
