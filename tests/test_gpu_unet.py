"""U-Net kernels (bf16 MFMA, FP32 accumulation) against PyTorch on the CPU in
FP64 with the same weights (MI355X).

Bars (bf16 activations and weights, FP32 accumulation): network output within
3e-2 relative L2 of the FP64 reference; every layer's weight and bias gradient
within 1e-1 (bf16 gradients through up to 14 layers); one Adam step moves each weight by lr * g / (|g| + eps) (the
first-step closed form) to 1e-3 relative.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONVS = [(16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 256), (256, 256),
         (384, 128), (128, 128), (192, 64), (64, 64), (96, 32), (32, 32)]


def _lib():
    from mantle_convection_lab_amd import hip_lib
    L = hip_lib()
    vp = C.c_void_p
    L.mcl_net_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)]
    L.mcl_net_destroy.argtypes = [vp]
    L.mcl_net_destroy.restype = None
    L.mcl_net_num_params.argtypes = [vp]
    L.mcl_net_num_params.restype = C.c_longlong
    for f in ("mcl_net_get_params", "mcl_net_set_params", "mcl_net_get_grads"):
        getattr(L, f).argtypes = [vp, C.POINTER(C.c_float)]
    L.mcl_net_forward.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.mcl_net_train_step.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                     C.POINTER(C.c_float)]
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _torch_net(params):
    import torch
    import torch.nn.functional as F
    ws, off = [], 0
    for cin, cout in CONVS:
        w = params[off: off + cout * 9 * cin].reshape(cout, 3, 3, cin)
        off += cout * 9 * cin
        b = params[off: off + cout]
        off += cout
        ws.append((torch.tensor(w, dtype=torch.float64).permute(0, 3, 1, 2).contiguous().requires_grad_(),
                   torch.tensor(b, dtype=torch.float64).requires_grad_()))
    wh = torch.tensor(params[off: off + 96].reshape(3, 32), dtype=torch.float64).requires_grad_()
    bh = torch.tensor(params[off + 96: off + 99], dtype=torch.float64).requires_grad_()

    def fwd(x):  # x: (B, 16, H, W)
        c = lambda i, t: F.relu(F.conv2d(t, ws[i][0], ws[i][1], padding=1))  # noqa: E731
        e1 = c(1, c(0, x))
        e2 = c(3, c(2, F.avg_pool2d(e1, 2)))
        e3 = c(5, c(4, F.avg_pool2d(e2, 2)))
        b = c(7, c(6, F.avg_pool2d(e3, 2)))
        d3 = c(9, c(8, torch.cat([F.interpolate(b, scale_factor=2, mode="nearest"), e3], 1)))
        d2 = c(11, c(10, torch.cat([F.interpolate(d3, scale_factor=2, mode="nearest"), e2], 1)))
        d1 = c(13, c(12, torch.cat([F.interpolate(d2, scale_factor=2, mode="nearest"), e1], 1)))
        return torch.einsum("oc,bchw->bohw", wh, d1) + bh[None, :, None, None]

    return fwd, ws, wh, bh


def _flat_grads(ws, wh, bh):
    out = []
    for w, b in ws:
        out.append(w.grad.permute(0, 2, 3, 1).reshape(-1).numpy())
        out.append(b.grad.numpy())
    out.append(wh.grad.reshape(-1).numpy())
    out.append(bh.grad.numpy())
    return out


def test_unet_forward_backward_adam_vs_torch():
    import torch
    L = _lib()
    H = W = 32
    B = 2
    net = C.c_void_p()
    assert L.mcl_net_create(H, W, B, 7, C.byref(net)) == 0
    try:
        n = L.mcl_net_num_params(net)
        assert n == sum(co * 9 * ci + co for ci, co in CONVS) + 99
        p = np.zeros(n, np.float32)
        assert L.mcl_net_get_params(net, _fp(p)) == 0
        rng = np.random.default_rng(3)
        x = np.zeros((B, H, W, 2), np.float32)
        x[..., 0] = rng.uniform(0, 1, (B, H, W))
        x[..., 1] = rng.uniform(-3, 0, (B, H, W))
        t = rng.normal(0, 1, (B, H, W, 3)).astype(np.float32)
        y = np.zeros((B, H, W, 3), np.float32)
        print("[gpu-unet] forward", flush=True)
        assert L.mcl_net_forward(net, _fp(x), _fp(y)) == 0
        fwd, ws, wh, bh = _torch_net(p.astype(np.float64))
        xt = torch.zeros((B, 16, H, W), dtype=torch.float64)
        xt[:, :2] = torch.tensor(x, dtype=torch.float64).permute(0, 3, 1, 2)
        yt = fwd(xt)
        yref = yt.detach().permute(0, 2, 3, 1).numpy()
        err = np.linalg.norm(y - yref) / np.linalg.norm(yref)
        print(f"  forward rel err {err:.2e}", flush=True)
        assert err < 3e-2
        print("[gpu-unet] gradients", flush=True)
        loss = C.c_float()
        assert L.mcl_net_train_step(net, _fp(x), _fp(t), 0.0, C.byref(loss)) == 0
        g = np.zeros(n, np.float32)
        assert L.mcl_net_get_grads(net, _fp(g)) == 0
        lt = ((yt - torch.tensor(t, dtype=torch.float64).permute(0, 3, 1, 2)) ** 2).mean()
        lt.backward()
        assert abs(loss.value - lt.item()) < 3e-2 * lt.item()
        ref = _flat_grads(ws, wh, bh)
        off = 0
        for i, r in enumerate(ref):
            gi = g[off: off + r.size]
            off += r.size
            e = np.linalg.norm(gi - r) / max(np.linalg.norm(r), 1e-30)
            print(f"  grad block {i}: rel err {e:.2e}", flush=True)
            assert e < 1e-1, i
        print("[gpu-unet] adam", flush=True)
        lr = 1e-3
        assert L.mcl_net_train_step(net, _fp(x), _fp(t), lr, C.byref(loss)) == 0
        p2 = np.zeros(n, np.float32)
        assert L.mcl_net_get_params(net, _fp(p2)) == 0
        expect = p - lr * g / (np.abs(g) + 1e-8)
        big = np.abs(g) > 1e-4
        assert np.max(np.abs(p2[big] - expect[big])) < 1e-3 * lr + 1e-7
    finally:
        L.mcl_net_destroy(net)


def test_staged_convolutions_match_direct_and_torch():
    """At 64 x 64 every 3x3 convolution runs the LDS-staged K = 32 MFMA kernel
    (64-pixel tiles of 1 x 64, 2 x 32, 4 x 16 and 8 x 8 pixels); the forward
    output and the gradients of one training step match the unstaged kernels
    (MCL_CONV_DIRECT=1) to bf16 rounding and PyTorch's FP64 result to the
    bars of the test above."""
    import os
    import torch
    L = _lib()
    H = W = 64
    B = 1
    rng = np.random.default_rng(11)
    x = np.zeros((B, H, W, 2), np.float32)
    x[..., 0] = rng.uniform(0, 1, (B, H, W))
    x[..., 1] = rng.uniform(-3, 0, (B, H, W))
    t = rng.normal(0, 1, (B, H, W, 3)).astype(np.float32)
    out = {}
    for mode in ("staged", "direct"):
        net = C.c_void_p()
        assert L.mcl_net_create(H, W, B, 7, C.byref(net)) == 0
        try:
            if mode == "direct":
                os.environ["MCL_CONV_DIRECT"] = "1"
            print(f"[gpu-unet] {mode} forward and gradients at 64 x 64", flush=True)
            n = L.mcl_net_num_params(net)
            p = np.zeros(n, np.float32)
            assert L.mcl_net_get_params(net, _fp(p)) == 0
            y = np.zeros((B, H, W, 3), np.float32)
            assert L.mcl_net_forward(net, _fp(x), _fp(y)) == 0
            loss = C.c_float()
            assert L.mcl_net_train_step(net, _fp(x), _fp(t), 0.0, C.byref(loss)) == 0
            g = np.zeros(n, np.float32)
            assert L.mcl_net_get_grads(net, _fp(g)) == 0
            out[mode] = (p, y, g)
        finally:
            os.environ.pop("MCL_CONV_DIRECT", None)
            L.mcl_net_destroy(net)
    p, ys, gs = out["staged"]
    _, yd, gd = out["direct"]
    assert np.array_equal(p, out["direct"][0])
    d = np.linalg.norm(ys - yd) / np.linalg.norm(yd)
    print(f"  staged vs direct forward: rel diff {d:.2e}", flush=True)
    assert d < 2e-2
    fwd, ws, wh, bh = _torch_net(p.astype(np.float64))
    xt = torch.zeros((B, 16, H, W), dtype=torch.float64)
    xt[:, :2] = torch.tensor(x, dtype=torch.float64).permute(0, 3, 1, 2)
    yt = fwd(xt)
    yref = yt.detach().permute(0, 2, 3, 1).numpy()
    err = np.linalg.norm(ys - yref) / np.linalg.norm(yref)
    print(f"  staged forward vs FP64: rel err {err:.2e}", flush=True)
    assert err < 3e-2
    lt = ((yt - torch.tensor(t, dtype=torch.float64).permute(0, 3, 1, 2)) ** 2).mean()
    lt.backward()
    off = 0
    for i, r in enumerate(_flat_grads(ws, wh, bh)):
        es = np.linalg.norm(gs[off: off + r.size] - r) / max(np.linalg.norm(r), 1e-30)
        ed = np.linalg.norm(gd[off: off + r.size] - r) / max(np.linalg.norm(r), 1e-30)
        off += r.size
        print(f"  grad block {i}: staged {es:.2e} direct {ed:.2e}", flush=True)
        assert es < 1e-1, i
