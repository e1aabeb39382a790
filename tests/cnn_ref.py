"""CPU (torch float32/float64) restatement of the learned preconditioner's network, used by
the tests as the CPU reference of the HIP kernels (qumasun_amd/csrc/qg_cnn.hip).

Parameter vector layout (2408 floats): W1 (8,2,27) | W2 (8,8,27) | W3 (1,8,27) | A1 B1 A2 B2 (8
each); tap t = 9 dz + 3 dy + dx is the offset (dz-1, dy-1, dx-1) in (z, y, x); grid arrays
are x-fastest, i = ix + Nx (iy + Ny iz), so a (Nz, Ny, Nx) view matches conv3d's (D, H, W).
"""
import numpy as np
import torch
import torch.nn.functional as F

C = 8
NPARAM = C * 2 * 27 + C * C * 27 + C * 27 + 4 * C
W8 = [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560]


def split(p):
    p = torch.as_tensor(p)
    o = 0
    out = {}
    for name, shape in (("W1", (C, 2, 3, 3, 3)), ("W2", (C, C, 3, 3, 3)), ("W3", (1, C, 3, 3, 3)),
                        ("A1", (C,)), ("B1", (C,)), ("A2", (C,)), ("B2", (C,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def conv(x, W):
    return F.conv3d(F.pad(x, (1,) * 6, mode="circular"), W)


def conv_adj(y, W):
    return F.conv3d(F.pad(y, (1,) * 6, mode="circular"), W.flip(2, 3, 4).transpose(0, 1))


def gates(P, v):
    g1 = 1 + torch.tanh(P["A1"].view(1, C, 1, 1, 1) * v + P["B1"].view(1, C, 1, 1, 1))
    g2 = 1 + torch.tanh(P["A2"].view(1, C, 1, 1, 1) * v + P["B2"].view(1, C, 1, 1, 1))
    return g1, g2


def bf16(x):
    """Round to bfloat16 (nearest even) and back: the matrix-core path's activation rounding."""
    return x.to(torch.bfloat16).to(x.dtype)


def net(P, v, u, rnd=lambda x: x):
    """N u for u of shape (B, 1, Nz, Ny, Nx); v of shape (1, 1, Nz, Ny, Nx).  rnd is applied to
    every convolution input (bf16 emulates the matrix-core inference kernels)."""
    g1, g2 = gates(P, v)
    x = torch.cat([u, v * u], 1)
    return conv(rnd(conv(rnd(conv(rnd(x), P["W1"]) * g1), P["W2"]) * g2), P["W3"])


def net_adj(P, v, y, rnd=lambda x: x):
    g1, g2 = gates(P, v)
    x = conv_adj(rnd(conv_adj(rnd(conv_adj(rnd(y), P["W3"]) * g2), P["W2"]) * g1), P["W1"])
    return x[:, :1] + v * x[:, 1:]


def kinetic_symbol(n, h):
    """T(k) of the periodic 8th-order -1/2 Laplacian on an (Nz, Ny, Nx) grid."""
    ts = []
    for d in range(3):
        k = np.fft.fftfreq(n[d]) * n[d]
        th = 2 * np.pi * k / n[d]
        ts.append(-0.5 * (W8[0] + 2 * sum(W8[m] * np.cos(m * th) for m in range(1, 5))) / h[d] ** 2)
    tx, ty, tz = ts
    return tz[:, None, None] + ty[None, :, None] + tx[None, None, :]


def kinetic(x, T, shift):
    Fx = torch.fft.fftn(x, dim=(2, 3, 4))
    return torch.fft.ifftn(Fx / (T + shift), dim=(2, 3, 4)).real


def loss(P, v, T, shift, u, t, znorm):
    """sum |K N^T N u - t|^2 / znorm."""
    y = kinetic(net_adj(P, v, net(P, v, u)), T, shift)
    return ((y - t) ** 2).sum() / znorm


def vnorm(V):
    V = np.asarray(V, np.float64)
    return (V - V.mean()) / V.std()
