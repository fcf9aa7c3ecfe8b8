"""Numpy restatement of the warm-start network of pffrg_mlp_* (test infrastructure).

Same layer layout (per layer W[out][in] row-major, then b[out]), SiLU hidden layers, linear
output, loss mean((y - t)^2).  With bf16=True every MFMA operand is rounded to BF16 (round to
nearest even) exactly where the kernels round it: the network input, the weights, the hidden
activations and the back-propagated dL/dz; products and sums are FP64.  With bf16=False it is
plain FP64 back-propagation, checked against finite differences."""
import numpy as np


def bf16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def dims_of(n_in, width, n_layers, n_out):
    return [n_in] + [width] * (n_layers - 1) + [n_out]


def unpack(params, dims):
    out, o = [], 0
    for l in range(len(dims) - 1):
        W = params[o:o + dims[l] * dims[l + 1]].reshape(dims[l + 1], dims[l])
        o += dims[l] * dims[l + 1]
        b = params[o:o + dims[l + 1]]
        o += dims[l + 1]
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(params, dims, x, bf=True):
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params, dims)
    h = r(x)
    hs, zs = [h], []
    for l, (W, b) in enumerate(layers):
        z = h @ r(W).T + b
        zs.append(z)
        if l + 1 < len(layers):
            h = r(z * _sig(z))
            hs.append(h)
    return zs[-1], hs, zs


def loss_grad(params, dims, x, t, bf=True):
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params, dims)
    y, hs, zs = forward(params, dims, x, bf)
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    dz = 2.0 * (y - t) / n
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dzr = r(dz)
        grads.append((dzr.T @ hs[l], dz.sum(axis=0)))
        if l > 0:
            dh = dzr @ r(W)
            s = _sig(zs[l - 1])
            dz = dh * (s * (1.0 + zs[l - 1] * (1.0 - s)))
    flat = []
    for dW, db in reversed(grads):
        flat.append(dW.reshape(-1))
        flat.append(db)
    return loss, np.concatenate(flat)
