"""CPU reference of the learned field step (test infrastructure).

The same Fourier neural operator as the library's g4mi_fno_*: lift 4 ->
width, Fourier layers z = IDFT(R . DFT(a)) + a W + b with GELU (tanh form)
after all but the last, projection width -> hidden (GELU) -> 2; truncated
DFTs over the modes k = -m/2 .. m/2-1 per axis written as explicit sums,
the inverse scaled by 1/N^2 and its real part taken.  Parameters come in
the library's flat order.  float64 torch on the CPU with autograd: the
gradient oracle for the GPU's hand-written backward pass.
"""
import math

import numpy as np


def blocks(N, width, modes, layers, hidden):
    C, H, nm2 = width, hidden, modes * modes
    out, o = {}, 0
    for name, n in [("P", 4 * C), ("bP", C)]:
        out[name] = (o, n)
        o += n
    for l in range(layers):
        for name, n in [(f"Rr{l}", nm2 * C * C), (f"Ri{l}", nm2 * C * C), (f"W{l}", C * C),
                        (f"b{l}", C)]:
            out[name] = (o, n)
            o += n
    for name, n in [("Q1", C * H), ("bQ1", H), ("Q2", 2 * H), ("bQ2", 2)]:
        out[name] = (o, n)
        o += n
    return out, o


def gelu(x):
    import torch
    c = math.sqrt(2.0 / math.pi)
    return 0.5 * x * (1.0 + torch.tanh(c * (x + 0.044715 * x ** 3)))


def forward(params, inp, N, width, modes, layers, hidden):
    """params: torch float64 vector; inp: [B, N*N, 4] -> out [B, N*N, 2]."""
    import torch
    C, H = width, hidden
    bl, _ = blocks(N, width, modes, layers, hidden)

    def blk(name, shape):
        o, n = bl[name]
        return params[o:o + n].reshape(shape)

    k = torch.arange(modes, dtype=torch.float64) - modes // 2
    x = torch.arange(N, dtype=torch.float64)
    th = 2 * math.pi * torch.outer(k, x) / N                 # [m, N]
    F = torch.exp(-1j * th.to(torch.complex128))             # forward kernel e^{-i k x}
    G = torch.exp(1j * th.to(torch.complex128))              # inverse kernel (1/N^2 below)
    B = inp.shape[0]
    a = inp @ blk("P", (4, C)) + blk("bP", (C,))
    for l in range(layers):
        a4 = a.reshape(B, N, N, C).to(torch.complex128)      # [b, y, x, c]
        ah = torch.einsum("kx,byxc->bykc", F, a4)            # x transform
        ah = torch.einsum("jy,bykc->bjkc", F, ah)            # y transform: [b, ky, kx, c]
        R = torch.complex(blk(f"Rr{l}", (modes, modes, C, C)), blk(f"Ri{l}", (modes, modes, C, C)))
        bh = torch.einsum("bjkc,jkco->bjko", ah, R)
        y1 = torch.einsum("jy,bjko->byko", G, bh)            # inverse y
        s = torch.einsum("kx,byko->byxo", G, y1).real / (N * N)  # inverse x, real part
        z = s.reshape(B, N * N, C) + a @ blk(f"W{l}", (C, C)) + blk(f"b{l}", (C,))
        a = gelu(z) if l < layers - 1 else z
    q = gelu(a @ blk("Q1", (C, H)) + blk("bQ1", (H,)))
    return q @ blk("Q2", (H, 2)) + blk("bQ2", (2,))


def loss_and_grad(params_np, inp_np, tgt_np, N, width, modes, layers, hidden):
    import torch
    p = torch.tensor(params_np.astype(np.float64), requires_grad=True)
    inp = torch.tensor(inp_np.astype(np.float64))
    tgt = torch.tensor(tgt_np.astype(np.float64))
    out = forward(p, inp, N, width, modes, layers, hidden)
    loss = torch.mean((out - tgt) ** 2)
    loss.backward()
    return float(loss.detach()), p.grad.numpy(), out.detach().numpy()
