"""torch.autograd.Function wrappers of the surrogate kernels (k_conv.hip).

Activations are NHWC bfloat16 tensors (n, h, w, c); weights and biases are
float32 masters, converted to bf16 operand layouts at every call.
"""

from __future__ import annotations

import torch

from .. import _lib


def _ctx(t):
    return _lib.ctx_for(t.device)


def _p(t):
    return _lib.ptr(t)


class _Conv3x3(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b, relu: bool, first8: bool = False):
        n, h, wd, cin = x.shape
        cout = w.shape[0]
        if w.shape != (cout, cin, 3, 3):
            raise ValueError(f"weight {tuple(w.shape)} does not match {cin} input channels")
        x = x.contiguous()
        w = w.contiguous()
        c = _ctx(x)
        wf = torch.empty(9 * cout * cin, dtype=torch.bfloat16, device=x.device)
        wt = torch.empty_like(wf)
        _lib.call("pff_conv_pack_weights", c, cin, cout, _p(w), _p(wf), _p(wt))
        y = torch.empty((n, h, wd, cout), dtype=torch.bfloat16, device=x.device)
        if first8:
            # channels 8.. of x are zero: the first-layer kernel (same bits as
            # the fused inference from the planes)
            import ctypes

            src = _lib.PffConvSrc(3, _p(x), None, None, 0, 0, None, None, None)
            _lib.call("pff_conv3x3_fwd_fused", c, n, h, wd, cin, cout, ctypes.byref(src), _p(wf), _p(b), int(relu),
                      _p(y))
        else:
            _lib.call("pff_conv3x3_fwd", c, n, h, wd, cin, cout, _p(x), _p(wf), _p(b), int(relu), _p(y))
        ctx.save_for_backward(x, y, wt)
        ctx.relu = relu
        ctx.shape = (n, h, wd, cin, cout)
        return y

    @staticmethod
    def backward(ctx, gy):
        x, y, wt = ctx.saved_tensors
        n, h, wd, cin, cout = ctx.shape
        c = _ctx(x)
        g = gy.to(torch.bfloat16).contiguous()
        if ctx.relu:
            g = g.clone()
            _lib.call("pff_relu_bwd", c, g.numel(), _p(y), _p(g))
        dx = None
        if ctx.needs_input_grad[0]:
            dx = torch.empty_like(x)
            _lib.call("pff_conv3x3_bwd_data", c, n, h, wd, cin, cout, _p(g), _p(wt), _p(dx))
        dw = torch.zeros((cout, cin, 3, 3), dtype=torch.float32, device=x.device)
        db = torch.zeros(cout, dtype=torch.float32, device=x.device)
        _lib.call("pff_conv3x3_bwd_weight", c, n, h, wd, cin, cout, _p(x), _p(g), _p(dw), _p(db))
        return dx, dw, db, None, None


class _AvgPool2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        n, h, w, cc = x.shape
        x = x.contiguous()
        y = torch.empty((n, h // 2, w // 2, cc), dtype=x.dtype, device=x.device)
        _lib.call("pff_avgpool2", _ctx(x), n, h, w, cc, _p(x), _p(y))
        ctx.shape = (n, h, w, cc)
        return y

    @staticmethod
    def backward(ctx, gy):
        n, h, w, cc = ctx.shape
        g = gy.to(torch.bfloat16).contiguous()
        dx = torch.empty((n, h, w, cc), dtype=torch.bfloat16, device=g.device)
        _lib.call("pff_avgpool2_bwd", _ctx(g), n, h, w, cc, _p(g), _p(dx))
        return dx


class _Upsample2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        n, h, w, cc = x.shape
        x = x.contiguous()
        y = torch.empty((n, 2 * h, 2 * w, cc), dtype=x.dtype, device=x.device)
        _lib.call("pff_upsample2", _ctx(x), n, h, w, cc, _p(x), _p(y))
        ctx.shape = (n, h, w, cc)
        return y

    @staticmethod
    def backward(ctx, gy):
        n, h, w, cc = ctx.shape
        g = gy.to(torch.bfloat16).contiguous()
        dx = torch.empty((n, h, w, cc), dtype=torch.bfloat16, device=g.device)
        _lib.call("pff_upsample2_bwd", _ctx(g), n, h, w, cc, _p(g), _p(dx))
        return dx


class _Concat(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b):
        n, h, w, ca = a.shape
        cb = b.shape[3]
        a = a.contiguous()
        b = b.contiguous()
        y = torch.empty((n, h, w, ca + cb), dtype=a.dtype, device=a.device)
        _lib.call("pff_concat", _ctx(a), n * h * w, ca, cb, _p(a), _p(b), _p(y))
        ctx.shape = (n, h, w, ca, cb)
        return y

    @staticmethod
    def backward(ctx, gy):
        n, h, w, ca, cb = ctx.shape
        g = gy.to(torch.bfloat16).contiguous()
        da = torch.empty((n, h, w, ca), dtype=torch.bfloat16, device=g.device)
        db = torch.empty((n, h, w, cb), dtype=torch.bfloat16, device=g.device)
        _lib.call("pff_concat_bwd", _ctx(g), n * h * w, ca, cb, _p(g), _p(da), _p(db))
        return da, db


class _Head1x1(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b):
        n, h, wd, cc = x.shape
        x = x.contiguous()
        y = torch.empty((n, h, wd), dtype=torch.float32, device=x.device)
        _lib.call("pff_head1x1", _ctx(x), n * h * wd, cc, _p(x), _p(w.contiguous()), _p(b), _p(y))
        ctx.save_for_backward(x, w)
        return y

    @staticmethod
    def backward(ctx, gy):
        x, w = ctx.saved_tensors
        n, h, wd, cc = x.shape
        g = gy.to(torch.float32).contiguous()
        dx = torch.empty_like(x) if ctx.needs_input_grad[0] else None
        dw = torch.zeros_like(w, dtype=torch.float32)
        db = torch.zeros(1, dtype=torch.float32, device=x.device)
        _lib.call("pff_head1x1_bwd", _ctx(x), n * h * wd, cc, _p(x), _p(w.contiguous()), _p(g), _p(dx), _p(dw),
                  _p(db))
        return dx, dw, db


def conv3x3(x, w, b, relu: bool = True, first8: bool = False):
    """3x3 convolution (stride 1, zero padding 1) of NHWC bf16 x with f32
    weights w[cout][cin][3][3] and bias b[cout]; optional fused ReLU.
    first8: x holds zeros past channel 7 (the packed input planes) and the
    first-layer kernel runs (cout 32, shapes tiled_ok; else the plain one)."""
    n, h, wd, _ = x.shape
    use8 = first8 and w.shape[0] == 32 and tiled_ok(h, wd)
    return _Conv3x3.apply(x, w, b, relu, use8)


def avgpool2(x):
    return _AvgPool2.apply(x)


def upsample2(x):
    return _Upsample2.apply(x)


def concat(a, b):
    return _Concat.apply(a, b)


def head1x1(x, w, b):
    """y[n, h, w] = b + sum_c x[..., c] w[c] in f32."""
    return _Head1x1.apply(x, w, b)


def pack_input(planes: torch.Tensor, channels: int = 32) -> torch.Tensor:
    """f32 planes (n, p, h, w) -> NHWC bf16 (n, h, w, channels), zero padded."""
    n, p, h, w = planes.shape
    out = torch.empty((n, h, w, channels), dtype=torch.bfloat16, device=planes.device)
    pl = planes.to(torch.float32).permute(1, 0, 2, 3).contiguous()  # [p][n*h*w]
    _lib.call("pff_pack_input", _ctx(pl), n * h * w, p, channels, _p(pl), _p(out))
    return out


def tiled_ok(h: int, w: int) -> bool:
    """Shapes the fused inference convolutions take (pff_conv3x3_fwd_fused)."""
    return w % 32 == 0 and h % 16 == 0


def conv3x3_fused(w, b, relu: bool = True, *, x=None, lo=None, skip=None, planes=None, head=None):
    """Inference-only 3x3 convolution that reads its input from the layers in
    front of it (no autograd): x (n, h, w, cin) NHWC bf16; or lo (n, h/2,
    w/2, ca) and skip (n, h, w, cb) for concat(upsample2(lo), skip); or planes
    (n, p, h, w) f32 zero padded to 32 channels (pack_input).  head = (hw, hb):
    returns the 1x1 head of the activated output, (n, h, w) f32, instead of
    the activations.  Same bits as the unfused ops."""
    cout, cin = w.shape[0], w.shape[1]
    if x is not None:
        n, h, wd, _ = x.shape
        src = _lib.PffConvSrc(0, _p(x.contiguous()), None, None, 0, 0, None, None, None)
        keep = (x,)
    elif lo is not None:
        n, h, wd, _ = skip.shape
        lo, skip = lo.contiguous(), skip.contiguous()
        src = _lib.PffConvSrc(1, _p(lo), _p(skip), None, lo.shape[3], 0, None, None, None)
        keep = (lo, skip)
    else:
        n, p, h, wd = planes.shape
        pl = planes.to(torch.float32).permute(1, 0, 2, 3).contiguous()  # [p][n*h*w]
        src = _lib.PffConvSrc(2, None, None, _p(pl), 0, p, None, None, None)
        keep = (pl,)
    dev = w.device
    wf = torch.empty(9 * cout * cin, dtype=torch.bfloat16, device=dev)
    wt = torch.empty_like(wf)
    c = _lib.ctx_for(dev)
    _lib.call("pff_conv_pack_weights", c, cin, cout, _p(w.contiguous()), _p(wf), _p(wt))
    y = None
    if head is not None:
        hw, hb = head[0].detach().contiguous(), head[1].detach().contiguous()
        out = torch.empty((n, h, wd), dtype=torch.float32, device=dev)
        src.head_w, src.head_b, src.head_out = _p(hw), _p(hb), _p(out)
        keep = keep + (hw, hb)
    else:
        y = torch.empty((n, h, wd, cout), dtype=torch.bfloat16, device=dev)
        out = y
    import ctypes

    _lib.call("pff_conv3x3_fwd_fused", c, n, h, wd, cin, cout, ctypes.byref(src), _p(wf), _p(b.detach()),
              int(relu), _p(y))
    del keep
    return out
