"""Neural reparametrisation of the velocity model (NNFWI), mirroring src/NN.jl:40-158 and
``add_initial_model`` (src/Utils.jl:140-166), in PyTorch.

The generator runs in torch on the GPU; its output velocity feeds the batched acoustic plan
through :func:`acoustic_traces`, whose pullback is the HIP adjoint, so one ``backward()``
gives the misfit gradient w.r.t. the network weights (examples/nn_fwi/NNFWI.jl:60-75).

Parity with the reference's Keras layers is not checkable here (no TensorFlow): the layer
sequence, sizes, padding and activations follow NN.jl, and the weights are freshly
initialised (Glorot-uniform, Keras' default).
"""
from __future__ import annotations

from typing import Optional

import torch
import torch.nn.functional as F


class Generator(torch.nn.Module):
    """Generator(z, isTrain; num_layer, h0, w0, vmin, vmax), src/NN.jl:40-96, and with
    ``dropout_rate`` the dropout variant of NN.jl:98-158.

    Dense(z -> h0*w0*8, no bias) -> reshape to (w0, h0, 8) -> tanh; then ``num_layer`` times
    bilinear 2x upsampling -> 4x4 convolution with 2^(num_layer-i+4) filters ("same" padding,
    no bias) -> LeakyReLU(0.1) [-> dropout]; a final 4x4 convolution to one channel; with
    vmin/vmax the output is ((vmax - vmin) tanh(x) + vmax + vmin) / 2.  ``forward(z)`` returns
    the squeezed image of shape (w0 * 2^num_layer, h0 * 2^num_layer) for one latent vector,
    (batch, ...) for several."""

    def __init__(self, z_size: int = 8, num_layer: int = 5, h0: int = 4, w0: int = 8,
                 vmin: Optional[float] = None, vmax: Optional[float] = None,
                 dropout_rate: Optional[float] = None):
        super().__init__()
        self.num_layer, self.h0, self.w0 = int(num_layer), int(h0), int(w0)
        self.vmin, self.vmax, self.dropout_rate = vmin, vmax, dropout_rate
        self.dense = torch.nn.Linear(z_size, self.h0 * self.w0 * 8, bias=False)
        convs, cin = [], 8
        for i in range(1, self.num_layer + 1):
            cout = 2 ** (self.num_layer - i + 4)
            convs.append(torch.nn.Conv2d(cin, cout, 4, bias=False))
            cin = cout
        self.convs = torch.nn.ModuleList(convs)
        self.out = torch.nn.Conv2d(cin, 1, 4, bias=False)
        for m in [self.dense, *self.convs, self.out]:
            torch.nn.init.xavier_uniform_(m.weight)

    @staticmethod
    def _same(x):
        # Keras "same" for a 4x4 kernel at stride 1: one row/column before, two after
        return F.pad(x, (1, 2, 1, 2))

    def forward(self, z: torch.Tensor) -> torch.Tensor:
        x = self.dense(z)
        # tf.reshape(x, [-1, w0, h0, 8]) is NHWC; keep its element order, then go NCHW
        x = x.reshape(-1, self.w0, self.h0, 8).permute(0, 3, 1, 2)
        x = torch.tanh(x)
        for conv in self.convs:
            x = F.interpolate(x, scale_factor=2, mode="bilinear", align_corners=False)
            x = F.leaky_relu(conv(self._same(x)), 0.1)
            if self.dropout_rate:
                x = F.dropout(x, self.dropout_rate, self.training)
        x = self.out(self._same(x))
        if self.vmin is not None and self.vmax is not None:
            x = ((self.vmax - self.vmin) * torch.tanh(x) + (self.vmax + self.vmin)) / 2.0
        return x.squeeze()


class UpsamplingGenerator(torch.nn.Module):
    """A 4-layer upsampling CNN velocity parametrisation: a fixed latent tensor of shape
    (latent_channels, h0, w0) drawn from a seeded standard normal; four times bilinear
    upsampling (the last one onto the output grid) -> 3x3 convolution ("same", with bias) with
    channels[0..3] outputs, LeakyReLU(0.1) after all but the last; velocity
    v = vmin + (vmax - vmin) * sigmoid(x).  ``forward()`` returns the (H, W) velocity image.
    The intermediate sizes double from (h0, w0) and the last upsampling lands on (H, W)."""

    def __init__(self, out_shape, latent_shape=(8, 16, 16), channels=(64, 32, 16, 1),
                 vmin: float = 1500.0, vmax: float = 4500.0, seed: int = 5,
                 dtype=torch.float32, device="cuda"):
        super().__init__()
        self.out_shape = tuple(int(s) for s in out_shape)
        self.vmin, self.vmax = float(vmin), float(vmax)
        g = torch.Generator().manual_seed(int(seed))
        z = torch.randn(1, *latent_shape, generator=g, dtype=torch.float64)
        self.register_buffer("z", z.to(dtype))
        convs, cin = [], int(latent_shape[0])
        for cout in channels:
            conv = torch.nn.Conv2d(cin, int(cout), 3, padding=1, bias=True)
            torch.nn.init.xavier_uniform_(conv.weight, generator=g)
            torch.nn.init.zeros_(conv.bias)
            convs.append(conv)
            cin = int(cout)
        self.convs = torch.nn.ModuleList(convs)
        h0, w0 = latent_shape[1], latent_shape[2]
        n = len(channels)
        self.sizes = [(h0 * 2 ** (k + 1), w0 * 2 ** (k + 1)) for k in range(n - 1)] + [self.out_shape]
        self.to(dtype=dtype, device=device)

    def forward(self) -> torch.Tensor:
        x = self.z
        for k, conv in enumerate(self.convs):
            x = F.interpolate(x, size=self.sizes[k], mode="bilinear", align_corners=False)
            x = conv(x)
            if k < len(self.convs) - 1:
                x = F.leaky_relu(x, 0.1)
        return self.vmin + (self.vmax - self.vmin) * torch.sigmoid(x[0, 0])


def add_initial_model(x: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
    """add_initial_model(x, v), src/Utils.jl:140-166: ``v`` plus the image ``x`` (or each of a
    batch of images), centred: a smaller image is added into the middle of ``v``, a larger one
    is cropped to ``v``'s size around its centre.  (The reference compares the size tuples
    lexicographically and fails on mixed cases; those raise here.)"""
    sx, sv = tuple(x.shape[-2:]), tuple(v.shape[-2:])
    if sx[0] <= sv[0] and sx[1] <= sv[1]:
        i0, j0 = (sv[0] - sx[0]) // 2, (sv[1] - sx[1]) // 2
        pad = (j0, sv[1] - sx[1] - j0, i0, sv[0] - sx[0] - i0)
        return v + F.pad(x, pad)
    if sx[0] >= sv[0] and sx[1] >= sv[1]:
        i0, j0 = (sx[0] - sv[0]) // 2, (sx[1] - sv[1]) // 2
        return v + x[..., i0:i0 + sv[0], j0:j0 + sv[1]]
    raise ValueError(f"Size error: {sv}, {sx}")
