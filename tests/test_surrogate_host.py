"""The warm-start U-Net's host side, without a GPU: the committed weights load
with the safe loader into the documented architecture, and the FLOP count the
benchmark's MFMA roofline divides by is what a counted forward performs."""

import math

import torch
import torch.nn.functional as F
from torch.utils.flop_counter import FlopCounterMode

from pyfastflow_amd.surrogate.unet import PLANES, UNet
from pyfastflow_amd.surrogate.warmstart import load_weights

# the documented channel plan: 4 levels of 32, 64, 128, 256 channels, two 3x3
# convolutions per level; the decoder convolves the upsampled map
# concatenated with the skip; 1x1 head from 32 channels to one
ENCODER = [(32, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 256), (256, 256)]
DECODER = [(256 + 128, 128), (128, 128), (128 + 64, 64), (64, 64), (64 + 32, 32), (32, 32)]


def test_committed_weights_load_into_the_documented_architecture():
    m = load_weights(UNet())  # torch.load(weights_only=True)
    assert m.shapes == ENCODER + DECODER
    n = sum(9 * ci * co + co for ci, co in ENCODER + DECODER) + 32 + 1
    assert sum(p.numel() for p in m.parameters()) == n == 1_955_233
    assert all(torch.isfinite(p).all() for p in m.parameters())
    lo, hi = (float(v) for v in m.target_range)
    assert math.isfinite(lo) and math.isfinite(hi) and lo < hi


def _torch_forward(m: UNet, planes: torch.Tensor) -> torch.Tensor:
    """The same network in plain torch ops (f32, CPU); the first convolution
    sees only the real input planes."""
    p = planes.shape[1]
    x = F.relu(F.conv2d(planes, m.w[0][:, :p], m.b[0], padding=1))
    x = F.relu(F.conv2d(x, m.w[1], m.b[1], padding=1))
    skips, i = [x], 2
    for _ in range(1, m.levels):
        x = F.avg_pool2d(x, 2)
        x = F.relu(F.conv2d(x, m.w[i], m.b[i], padding=1))
        x = F.relu(F.conv2d(x, m.w[i + 1], m.b[i + 1], padding=1))
        skips.append(x)
        i += 2
    for k in range(m.levels - 2, -1, -1):
        x = torch.cat([F.interpolate(x, scale_factor=2, mode="nearest"), skips[k]], dim=1)
        x = F.relu(F.conv2d(x, m.w[i], m.b[i], padding=1))
        x = F.relu(F.conv2d(x, m.w[i + 1], m.b[i + 1], padding=1))
        i += 2
    return F.conv2d(x, m.hw.view(1, -1, 1, 1), m.hb)[:, 0]


def test_flops_per_pixel_is_what_a_counted_forward_performs():
    m = UNet()
    h = w = 64
    planes = torch.randn(1, len(PLANES), h, w)
    with torch.no_grad(), FlopCounterMode(display=False) as fc:
        y = _torch_forward(m, planes)
    assert y.shape == (1, h, w)
    assert fc.get_total_flops() == m.flops_per_pixel() * h * w
    # 0.33 MFLOP per pixel: 1.37 TFLOP for a 2048 x 2048 inference
    assert abs(m.flops_per_pixel() * 2048 * 2048 / 1e12 - 1.368) < 1e-3
