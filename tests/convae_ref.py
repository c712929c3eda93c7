"""PyTorch (CPU, FP32) reference of the convolutional autoencoder snapshot codec.

The network: 3x3x3 convolutions with zero padding 1, channels 1->16 (stride 2), 16->32 (stride 2),
32->32, 32->4 (latent), 4->32, nearest x2 + 32->32, nearest x2 + 32->16, 16->1, exact (erf) GELU after
every layer except the latent and the output.  Blocks are channels-last [nb][D^3][C] with position
(iy*D + ix)*D + iz on the device; here they are [nb][C][D][D][D] with the spatial axes (y, x, z), and the
weights [cout][cin][27] with tap (dy+1)*9 + (dx+1)*3 + dz+1 are [cout][cin][3][3][3] -- torch's conv3d is
the same cross-correlation.  `bf16=True` rounds every convolution's operands (input and weights) to
BF16 as the matrix-core kernels do, so the reference differs from the device only by the FP32
summation order."""
import numpy as np
import torch
import torch.nn.functional as F

CIN = [1, 16, 32, 32, 4, 32, 32, 16]
COUT = [16, 32, 32, 4, 32, 32, 16, 1]
STRIDE = [2, 2, 1, 1, 1, 1, 1, 1]
UP = [False, False, False, False, False, True, True, False]
GELU = [True, True, True, False, True, True, True, False]


def split_params(p):
    """Flat parameter vector -> [(W [cout][cin][3][3][3], b [cout])] per layer."""
    out, o = [], 0
    for ci, co in zip(CIN, COUT):
        w = p[o:o + co * ci * 27].reshape(co, ci, 3, 3, 3)
        o += co * ci * 27
        b = p[o:o + co]
        o += co
        out.append((w, b))
    assert o == p.size
    return out


def _bf(t, on):
    return t.to(torch.bfloat16).to(torch.float32) if on else t


def layer(l, x, w, b, bf16):
    if UP[l]:
        x = F.interpolate(x, scale_factor=2, mode="nearest")
    z = F.conv3d(_bf(x, bf16), _bf(w, bf16), b, stride=STRIDE[l], padding=1)
    return z, (F.gelu(z) if GELU[l] else z)


def forward(params, x, bf16=True, l0=0, l1=8):
    """x: [nb][C][D][D][D] torch tensor -> output of layer l1-1 (and the latent when it is passed)."""
    layers = split_params(params)
    a = x
    latent = None
    for l in range(l0, l1):
        w, b = (torch.from_numpy(np.ascontiguousarray(t)) for t in layers[l])
        _, a = layer(l, a, w, b, bf16)
        if l == 3:
            latent = a
    return a, latent


def loss_and_grad(params, x_np):
    """Mean squared reconstruction error of blocks x [nb][32^3] and its gradient (FP32, no rounding)."""
    p = torch.from_numpy(params.astype(np.float32)).requires_grad_(True)
    x = torch.from_numpy(x_np.reshape(-1, 1, 32, 32, 32))
    ws, o = [], 0
    a = x
    for l, (ci, co) in enumerate(zip(CIN, COUT)):
        w = p[o:o + co * ci * 27].reshape(co, ci, 3, 3, 3)
        o += co * ci * 27
        b = p[o:o + co]
        o += co
        _, a = layer(l, a, w, b, False)
    loss = ((a - x) ** 2).mean()
    loss.backward()
    return float(loss.detach()), p.grad.numpy().copy()


def to_torch(blocks, C, D):
    """channels-last device layout [nb][D^3][C] -> [nb][C][D][D][D]"""
    return torch.from_numpy(np.ascontiguousarray(blocks.reshape(-1, D, D, D, C).transpose(0, 4, 1, 2, 3)))


def to_device_layout(t):
    """[nb][C][D][D][D] -> [nb][D^3][C]"""
    a = t.detach().numpy()
    nb, C, D = a.shape[0], a.shape[1], a.shape[2]
    return np.ascontiguousarray(a.transpose(0, 2, 3, 4, 1).reshape(nb, D ** 3, C))


def gather_blocks(vol, floor_frac):
    """The codec's blocks of a volume [ny][nx][nz]: zero-padded 32^3 blocks in order (by, bx, bz),
    each divided by its scale rms(block) + floor_frac*rms(volume)."""
    ny, nx, nz = vol.shape
    nby, nbx, nbz = (ny + 31) // 32, (nx + 31) // 32, (nz + 31) // 32
    pad = np.zeros((nby * 32, nbx * 32, nbz * 32), np.float32)
    pad[:ny, :nx, :nz] = vol
    rms_v = np.sqrt(np.mean(vol.astype(np.float64) ** 2))
    blocks, scales = [], []
    for by in range(nby):
        for bx in range(nbx):
            for bz in range(nbz):
                blk = pad[by * 32:(by + 1) * 32, bx * 32:(bx + 1) * 32, bz * 32:(bz + 1) * 32]
                s = np.sqrt(np.mean(blk.astype(np.float64) ** 2)) + floor_frac * rms_v
                blocks.append(blk / s if s > 0 else blk * 0)
                scales.append(s)
    return np.array(blocks, np.float32), np.array(scales)


def scatter_blocks(blocks, scales, shape):
    ny, nx, nz = shape
    nby, nbx, nbz = (ny + 31) // 32, (nx + 31) // 32, (nz + 31) // 32
    pad = np.zeros((nby * 32, nbx * 32, nbz * 32), np.float32)
    i = 0
    for by in range(nby):
        for bx in range(nbx):
            for bz in range(nbz):
                pad[by * 32:(by + 1) * 32, bx * 32:(bx + 1) * 32, bz * 32:(bz + 1) * 32] = blocks[i] * scales[i]
                i += 1
    return pad[:ny, :nx, :nz]
