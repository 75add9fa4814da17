"""Plain PyTorch FP32 (CPU) restatement of the denoiser network, forward and
backward, rounding to BF16 at the same points as the device kernels: stored
activations, the loss gradient and every stored input gradient.  Test
infrastructure only."""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

C = 32
TAPS = 27
LAYER_W = C * TAPS * C
LAYER_P = LAYER_W + C
LAYERS = 18
BLOCKS = 8


def bf(t: torch.Tensor) -> torch.Tensor:
    return t.to(torch.bfloat16).to(t.dtype)


def split(params: np.ndarray):
    """[(W (co, ci, 3, 3, 3) torch layout, b)] per layer from the flat array
    ([co][tap][ci] weights then 32 biases per layer, tap = kz*9 + ky*3 + kx)."""
    out = []
    for l in range(LAYERS):
        p = params[l * LAYER_P:(l + 1) * LAYER_P]
        w = torch.from_numpy(p[:LAYER_W].reshape(C, 3, 3, 3, C).transpose(0, 4, 1, 2, 3).copy())
        b = torch.from_numpy(p[LAYER_W:].copy())
        out.append((w, b))
    return out


def flat_grad(dw: torch.Tensor, db: torch.Tensor) -> np.ndarray:
    return np.concatenate([dw.permute(0, 2, 3, 4, 1).reshape(-1).numpy(), db.numpy()]).astype(np.float64)


def conv(x, w, b=None):
    """x channel-last (N, D, H, W, C)."""
    y = F.conv3d(x.permute(0, 4, 1, 2, 3), w, b, padding=1)
    return y.permute(0, 2, 3, 4, 1)


def conv_input_grad(g, w):
    shape = (g.shape[0], C, g.shape[1], g.shape[2], g.shape[3])
    r = torch.nn.grad.conv3d_input(shape, w, g.permute(0, 4, 1, 2, 3), padding=1)
    return r.permute(0, 2, 3, 4, 1)


def conv_weight_grad(x, g):
    return torch.nn.grad.conv3d_weight(x.permute(0, 4, 1, 2, 3), (C, C, 3, 3, 3), g.permute(0, 4, 1, 2, 3), padding=1)


def pad_features(f: np.ndarray, dtype=torch.float32) -> torch.Tensor:
    x = torch.zeros(f.shape[:-1] + (C,), dtype=dtype)
    x[..., :f.shape[-1]] = torch.from_numpy(f).to(dtype)
    return x


def _identity(t: torch.Tensor) -> torch.Tensor:
    return t


def forward(params: np.ndarray, f: np.ndarray, rounding: bool = True, dtype=torch.float32):
    """Returns (pred (N, D, H, W), saved tensors).  rounding=False gives the
    unrounded network (used to check the backward against autograd)."""
    rnd = bf if rounding else _identity
    layers = [(rnd(w.to(dtype)), b.to(dtype)) for w, b in split(params)]
    x0 = rnd(pad_features(f, dtype))
    a = [rnd(torch.relu(conv(x0, *layers[0])))]
    h = []
    for k in range(BLOCKS):
        h.append(rnd(torch.relu(conv(a[k], *layers[2 * k + 1]))))
        a.append(rnd(torch.relu(conv(h[k], *layers[2 * k + 2]) + a[k])))
    out = conv(a[BLOCKS], *layers[LAYERS - 1])[..., 0]
    pred = torch.from_numpy(f[..., 0]).to(dtype) + out
    return pred, (x0, a, h, layers)


def loss_and_grads(params: np.ndarray, f: np.ndarray, target: np.ndarray, mask: np.ndarray, rounding: bool = True,
                   dtype=torch.float32):
    rnd = bf if rounding else _identity
    pred, (x0, a, h, layers) = forward(params, f, rounding, dtype)
    t = torch.from_numpy(target).to(dtype)
    m = torch.from_numpy(mask).to(dtype)
    count = float(mask.astype(np.float64).sum())
    diff = pred - t
    loss = float((m * diff * diff).sum()) / count
    gout = torch.zeros(pred.shape + (C,), dtype=dtype)
    gout[..., 0] = rnd(2.0 * m * diff * (1.0 / count))
    grads = [None] * LAYERS
    grads[LAYERS - 1] = flat_grad(conv_weight_grad(a[BLOCKS], gout), gout.sum(dim=(0, 1, 2, 3)))
    gs = rnd(conv_input_grad(gout, layers[LAYERS - 1][0]) * (a[BLOCKS] > 0))
    for k in range(BLOCKS - 1, -1, -1):
        grads[2 * k + 2] = flat_grad(conv_weight_grad(h[k], gs), gs.sum(dim=(0, 1, 2, 3)))
        gh = rnd(conv_input_grad(gs, layers[2 * k + 2][0]) * (h[k] > 0))
        grads[2 * k + 1] = flat_grad(conv_weight_grad(a[k], gh), gh.sum(dim=(0, 1, 2, 3)))
        gs = rnd((gs + conv_input_grad(gh, layers[2 * k + 1][0])) * (a[k] > 0))
    grads[0] = flat_grad(conv_weight_grad(x0, gs), gs.sum(dim=(0, 1, 2, 3)))
    return loss, pred.numpy(), np.concatenate(grads)


def random_params(seed: int, scale: float = 1.0) -> np.ndarray:
    """He-scaled random network with the padded channels at zero and a
    non-zero correction layer (so every layer receives a gradient)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(LAYERS * LAYER_P, dtype=np.float32)
    for l in range(LAYERS):
        ci = 4 if l == 0 else C
        co = 1 if l == LAYERS - 1 else C
        w = np.zeros((C, TAPS, C), dtype=np.float32)
        w[:co, :, :ci] = rng.normal(0.0, scale * np.sqrt(2.0 / (TAPS * ci)), (co, TAPS, ci))
        p[l * LAYER_P:l * LAYER_P + LAYER_W] = w.reshape(-1)
        p[l * LAYER_P + LAYER_W:l * LAYER_P + LAYER_W + co] = rng.normal(0.0, 0.05, co)
    return p
