"""The warm-start U-Net: 4 levels (32, 64, 128, 256 channels), two 3x3
convolutions with ReLU per level, 2x2 average pooling down, nearest 2x
upsampling and skip concatenation up, and a 1x1 head to log water depth."""

from __future__ import annotations

import math

import torch
from torch import nn

from . import ops

PLANES = ("z", "log_area", "log_slope", "log_fill_depth", "rain", "manning")


class UNet(nn.Module):
    def __init__(self, base: int = 32, levels: int = 4, in_planes: int = len(PLANES), seed: int = 0):
        super().__init__()
        self.base, self.levels, self.in_planes = base, levels, in_planes
        g = torch.Generator().manual_seed(seed)
        self.w = nn.ParameterList()
        self.b = nn.ParameterList()
        self.shapes = []

        def add(cin, cout):
            std = math.sqrt(2.0 / (9 * cin))
            self.w.append(nn.Parameter(torch.randn(cout, cin, 3, 3, generator=g) * std))
            self.b.append(nn.Parameter(torch.zeros(cout)))
            self.shapes.append((cin, cout))

        ch = [base * 2**k for k in range(levels)]
        cin = base  # the input planes are zero padded to `base` channels
        for k in range(levels):
            add(cin, ch[k])
            add(ch[k], ch[k])
            cin = ch[k]
        for k in range(levels - 2, -1, -1):
            add(ch[k + 1] + ch[k], ch[k])
            add(ch[k], ch[k])
        self.hw = nn.Parameter(torch.randn(base, generator=g) * math.sqrt(1.0 / base))
        self.hb = nn.Parameter(torch.zeros(1))
        # range of the training targets (log depth); predictions are clamped to it
        self.register_buffer("target_range", torch.tensor([-10.0, 10.0]))

    def forward(self, planes: torch.Tensor) -> torch.Tensor:
        """planes (n, in_planes, h, w) f32, h and w multiples of 2^(levels-1)
        -> predicted log water depth (n, h, w) f32.  Without gradients and on
        shapes the fused kernels take, the input packing, the upsampling and
        concatenation and the head run inside the convolutions (same bits)."""
        if not torch.is_grad_enabled() and self._fused_ok(planes):
            return self._forward_fused(planes)
        x = ops.pack_input(planes, self.base)
        skips = []
        i = 0
        for k in range(self.levels):
            if k > 0:
                x = ops.avgpool2(x)
            # the first convolution sees at most 8 nonzero channels when the
            # planes are few: the first-layer kernel, as the fused inference
            x = ops.conv3x3(x, self.w[i], self.b[i], first8=(i == 0 and planes.shape[1] <= 8))
            x = ops.conv3x3(x, self.w[i + 1], self.b[i + 1])
            i += 2
            skips.append(x)
        for k in range(self.levels - 2, -1, -1):
            x = ops.concat(ops.upsample2(x), skips[k])
            x = ops.conv3x3(x, self.w[i], self.b[i])
            x = ops.conv3x3(x, self.w[i + 1], self.b[i + 1])
            i += 2
        return ops.head1x1(x, self.hw, self.hb)

    def _fused_ok(self, planes: torch.Tensor) -> bool:
        _, p, h, w = planes.shape
        if p > self.base or self.base != 32:
            return False
        return all(ops.tiled_ok(h >> k, w >> k) for k in range(self.levels))

    def _forward_fused(self, planes: torch.Tensor) -> torch.Tensor:
        w, b = self.w, self.b
        x = ops.conv3x3_fused(w[0], b[0], planes=planes)
        x = ops.conv3x3_fused(w[1], b[1], x=x)
        skips = [x]
        i = 2
        for k in range(1, self.levels):
            x = ops.avgpool2(x)
            x = ops.conv3x3_fused(w[i], b[i], x=x)
            x = ops.conv3x3_fused(w[i + 1], b[i + 1], x=x)
            i += 2
            skips.append(x)
        for k in range(self.levels - 2, -1, -1):
            x = ops.conv3x3_fused(w[i], b[i], lo=x, skip=skips[k])
            last = k == 0
            x = ops.conv3x3_fused(w[i + 1], b[i + 1], x=x, head=(self.hw, self.hb) if last else None)
            i += 2
        return x

    def flops_per_pixel(self) -> float:
        """Multiply-adds x 2 of the 3x3 convolutions and the head, per input
        pixel; the first convolution counts its real input planes, not the
        zero channels they are padded with."""
        f = 0.0
        i = 0
        for k in range(self.levels):
            scale = 4.0 ** (-k)
            for _ in range(2):
                cin, cout = self.shapes[i]
                if i == 0:
                    cin = min(cin, self.in_planes)
                f += 2 * 9 * cin * cout * scale
                i += 1
        for k in range(self.levels - 2, -1, -1):
            scale = 4.0 ** (-k)
            for _ in range(2):
                cin, cout = self.shapes[i]
                f += 2 * 9 * cin * cout * scale
                i += 1
        return f + 2 * self.base
