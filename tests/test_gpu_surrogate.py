"""GPU checks of the surrogate kernels against torch on the CPU (float64 on the
same bf16-rounded operands).

Forward and data-gradient outputs are bf16: each element must be within one
bf16 rounding (2^-8 relative) of the float64 result, plus 1e-3 of the
tensor's largest magnitude for accumulation-order effects.  Weight gradients
are f32 sums over every pixel: within 1e-3 of the largest magnitude.
"""

import numpy as np
import pytest
import torch
import torch.nn.functional as F

pytestmark = pytest.mark.gpu


def _bf(t):
    return t.to(torch.bfloat16)


def _ref64(t):
    return t.detach().to("cpu", torch.float64)


def _close(got, ref, rel=2.0**-8, absf=1e-3, what=""):
    got = got.detach().to("cpu", torch.float64)
    ref = ref.detach().to("cpu", torch.float64)
    tol = rel * ref.abs() + absf * float(ref.abs().max()) + 1e-12
    err = (got - ref).abs()
    bad = err > tol
    assert not bool(bad.any()), f"{what}: {int(bad.sum())} of {bad.numel()} out of tolerance, max err {float(err.max())}"


@pytest.mark.parametrize("n,h,w,cin,cout,relu", [
    (1, 5, 7, 32, 32, False),
    (2, 16, 70, 32, 64, True),
    (1, 33, 64, 64, 32, False),
    (1, 8, 130, 96, 128, True),
    # tiled kernel (W a multiple of 32, H of 16): 16 x 32 and 8 x 32 tiles
    (1, 16, 128, 32, 32, False),
    (2, 32, 96, 64, 64, True),
    (1, 16, 64, 64, 128, False),
    (3, 16, 32, 96, 32, True),
    (1, 48, 160, 160, 64, False),
])
def test_conv3x3_forward(pf, n, h, w, cin, cout, relu):
    from pyfastflow_amd.surrogate import ops

    print(f"case conv fwd {n}x{h}x{w} {cin}->{cout} relu={relu}", flush=True)
    g = torch.Generator().manual_seed(n * 1000 + h * 10 + w)
    x = _bf(torch.randn(n, h, w, cin, generator=g))
    wt = torch.randn(cout, cin, 3, 3, generator=g) * 0.1
    b = torch.randn(cout, generator=g) * 0.1
    y = ops.conv3x3(x.cuda(), wt.cuda(), b.cuda(), relu=relu)
    xr = _ref64(x).permute(0, 3, 1, 2)
    wr = _ref64(_bf(wt))
    ref = F.conv2d(xr, wr, _ref64(b), padding=1).permute(0, 2, 3, 1)
    if relu:
        ref = ref.clamp_min(0)
    _close(y, ref, what="conv3x3 forward")


@pytest.mark.parametrize("n,h,w,cin,cout", [(1, 6, 9, 32, 32), (2, 16, 66, 64, 32), (1, 12, 40, 32, 64),
                                             (2, 16, 128, 32, 64), (1, 32, 32, 64, 32), (2, 8, 64, 96, 32),
                                             (1, 5, 256, 32, 32), (3, 4, 32, 256, 128), (2, 3, 96, 64, 64)])
def test_conv3x3_backward(pf, n, h, w, cin, cout):
    from pyfastflow_amd.surrogate import ops

    print(f"case conv bwd {n}x{h}x{w} {cin}->{cout}", flush=True)
    g = torch.Generator().manual_seed(7 + h + w)
    x = _bf(torch.randn(n, h, w, cin, generator=g))
    wt = torch.randn(cout, cin, 3, 3, generator=g) * 0.1
    b = torch.randn(cout, generator=g) * 0.1
    gy = _bf(torch.randn(n, h, w, cout, generator=g))
    xg = x.cuda().requires_grad_(True)
    wg = wt.cuda().requires_grad_(True)
    bg = b.cuda().requires_grad_(True)
    y = ops.conv3x3(xg, wg, bg, relu=True)
    y.backward(gy.cuda())
    # float64 reference on the same bf16 operands; the ReLU mask is taken from
    # the kernel's own (bf16) output so that ties at 0 agree
    mask = (y.detach().cpu().to(torch.float64) > 0).permute(0, 3, 1, 2)
    xr = _ref64(x).permute(0, 3, 1, 2).requires_grad_(True)
    wr = _ref64(_bf(wt)).requires_grad_(True)
    br = _ref64(b).requires_grad_(True)
    yr = F.conv2d(xr, wr, br, padding=1)
    yr.backward(_ref64(gy).permute(0, 3, 1, 2) * mask)
    _close(xg.grad, xr.grad.permute(0, 2, 3, 1), what="dx")
    _close(wg.grad, wr.grad, rel=1e-3, absf=1e-3, what="dw")
    _close(bg.grad, br.grad, rel=1e-3, absf=1e-3, what="db")


@pytest.mark.parametrize("n,h,w,p", [(1, 16, 32, 6), (2, 16, 64, 1), (1, 32, 96, 8), (3, 48, 32, 3)])
def test_conv3x3_first_layer(pf, n, h, w, p):
    """The first-layer kernel (k = tap x 8 channels; x zero past channel 7)
    against float64, forward and gradients, and the fused form that reads
    the f32 planes gives the same bits."""
    from pyfastflow_amd.surrogate import ops

    print(f"case conv first {n}x{h}x{w} planes={p}", flush=True)
    g = torch.Generator().manual_seed(100 + p + h)
    planes = torch.randn(n, p, h, w, generator=g)
    x = ops.pack_input(planes.cuda(), 32).cpu()
    assert torch.count_nonzero(x[..., p:]) == 0
    wt = torch.randn(32, 32, 3, 3, generator=g) * 0.1
    b = torch.randn(32, generator=g) * 0.1
    gy = _bf(torch.randn(n, h, w, 32, generator=g))
    xg = x.cuda()
    wg = wt.cuda().requires_grad_(True)
    bg = b.cuda().requires_grad_(True)
    y = ops.conv3x3(xg, wg, bg, relu=True, first8=True)
    y.backward(gy.cuda())
    mask = (y.detach().cpu().to(torch.float64) > 0).permute(0, 3, 1, 2)
    xr = _ref64(x).permute(0, 3, 1, 2)
    wr = _ref64(_bf(wt)).requires_grad_(True)
    br = _ref64(b).requires_grad_(True)
    yr = F.conv2d(xr, wr, br, padding=1)
    _close(y, yr.clamp_min(0).permute(0, 2, 3, 1), what="first-layer forward")
    yr.backward(_ref64(gy).permute(0, 3, 1, 2) * mask)
    _close(wg.grad, wr.grad, rel=1e-3, absf=1e-3, what="first-layer dw")
    _close(bg.grad, br.grad, rel=1e-3, absf=1e-3, what="first-layer db")
    with torch.no_grad():
        yf = ops.conv3x3_fused(wt.cuda(), b.cuda(), planes=planes.cuda())
    assert torch.equal(yf.view(torch.int16).cpu(), y.detach().view(torch.int16).cpu())


def test_pool_upsample_concat_head(pf):
    from pyfastflow_amd.surrogate import ops

    g = torch.Generator().manual_seed(3)
    x = _bf(torch.randn(2, 8, 12, 32, generator=g))
    xc = x.cuda().requires_grad_(True)
    p = ops.avgpool2(xc)
    pr = F.avg_pool2d(_ref64(x).permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1)
    _close(p, pr, what="avgpool2")
    u = ops.upsample2(p)
    ur = F.interpolate(pr.permute(0, 3, 1, 2), scale_factor=2, mode="nearest").permute(0, 2, 3, 1)
    _close(u, ur, what="upsample2")
    c = ops.concat(u, xc)
    assert torch.equal(c[..., :32].cpu(), u.cpu()) and torch.equal(c[..., 32:].cpu(), xc.detach().cpu())
    hw = torch.randn(64, generator=g).cuda().requires_grad_(True)
    hb = torch.randn(1, generator=g).cuda().requires_grad_(True)
    y = ops.head1x1(c, hw, hb)
    yr = (_ref64(c) * _ref64(hw)).sum(-1) + float(hb.detach())
    _close(y, yr, rel=1e-5, absf=1e-5, what="head")
    gy = torch.randn(y.shape, generator=g)
    y.backward(gy.cuda())
    # reference gradient through the same chain in float64
    x64 = _ref64(x).requires_grad_(True)
    hw64 = _ref64(hw).requires_grad_(True)
    p64 = F.avg_pool2d(x64.permute(0, 3, 1, 2), 2)
    u64 = F.interpolate(p64, scale_factor=2, mode="nearest").permute(0, 2, 3, 1)
    c64 = torch.cat([u64, x64], -1)
    y64 = (c64 * hw64).sum(-1)
    y64.backward(_ref64(gy))
    _close(xc.grad, x64.grad, rel=2.0**-7, absf=2e-3, what="dx through pool/upsample/concat/head")
    # the head's weight gradient sums dy * (its bf16 input), so take the
    # kernel's own concatenated input as the operand
    hw_ref = (_ref64(c) * _ref64(gy)[..., None]).sum((0, 1, 2))
    _close(hw.grad, hw_ref, rel=1e-4, absf=1e-4, what="head dw")


def test_unet_gradients_against_cpu(pf):
    """The whole U-Net: loss gradients from the HIP kernels against the same
    network in float64 on the CPU with bf16 rounding of the weights (the
    activations stay f64 there, so the bar is on direction and size)."""
    from pyfastflow_amd.surrogate import UNet

    torch.manual_seed(0)
    m = UNet(seed=1).cuda()
    planes = torch.randn(2, 4, 32, 32)
    target = torch.randn(2, 32, 32)
    out = m(planes.cuda())
    loss = ((out - target.cuda()) ** 2).mean()
    loss.backward()

    def cpu_forward(params, planes64):
        w = [p for p in params[:-2]]
        ws, bs = w[0::2], w[1::2]
        x = torch.zeros(planes64.shape[0], m.base, *planes64.shape[2:], dtype=torch.float64)
        x[:, : planes64.shape[1]] = planes64
        skips, i = [], 0
        for k in range(m.levels):
            if k > 0:
                x = F.avg_pool2d(x, 2)
            for _ in range(2):
                x = F.relu(F.conv2d(x, ws[i].to(torch.bfloat16).to(torch.float64), bs[i], padding=1))
                i += 1
            skips.append(x)
        for k in range(m.levels - 2, -1, -1):
            x = torch.cat([F.interpolate(x, scale_factor=2, mode="nearest"), skips[k]], 1)
            for _ in range(2):
                x = F.relu(F.conv2d(x, ws[i].to(torch.bfloat16).to(torch.float64), bs[i], padding=1))
                i += 1
        return torch.einsum("nchw,c->nhw", x, params[-2]) + params[-1]

    order = []
    for wi, bi in zip(m.w, m.b):
        order += [wi, bi]
    order += [m.hw, m.hb]
    p64 = [p.detach().cpu().to(torch.float64).requires_grad_(True) for p in order]
    o64 = cpu_forward(p64, planes.to(torch.float64))
    l64 = ((o64 - target.to(torch.float64)) ** 2).mean()
    l64.backward()
    lv, rv = float(loss.detach()), float(l64.detach())
    assert abs(lv - rv) <= 0.02 * abs(rv) + 1e-3
    for k, (pg, pr) in enumerate(zip(order, p64)):
        a = pg.grad.detach().cpu().to(torch.float64).ravel()
        r = pr.grad.ravel()
        cos = float(torch.dot(a, r) / (a.norm() * r.norm() + 1e-30))
        rel = float((a - r).norm() / (r.norm() + 1e-30))
        assert cos > 0.99 and rel < 0.1, f"parameter {k}: cosine {cos:.4f}, relative error {rel:.4f}"


@pytest.mark.parametrize("size", [256, 512])
def test_unet_fused_inference_same_bits(pf, size):
    """Inference with the fused convolutions (input packing, upsampling +
    concatenation and head read inside the convolutions) gives the same bits
    as the layer-by-layer forward pass."""
    from pyfastflow_amd.surrogate import UNet

    m = UNet(seed=3).cuda()
    g = torch.Generator().manual_seed(size)
    planes = torch.randn(1, m.in_planes, size, size, generator=g).cuda()
    with torch.no_grad():
        assert m._fused_ok(planes)
        fused = m(planes)
    ref = m(planes.requires_grad_(True)).detach()  # grad enabled: the unfused path
    assert fused.shape == ref.shape
    assert torch.equal(fused.view(torch.int32), ref.view(torch.int32))
