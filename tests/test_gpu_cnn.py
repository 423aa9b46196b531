"""The learned correction's kernels against a CPU FP64 reference.

The reference is torch on the CPU in float64 (conv3d / conv_transpose3d /
conv3d_weight), an implementation independent of the HIP kernels.  It is
evaluated twice: with the kernels' bf16 rounding points emulated (operands
of the 32-channel layers, activations, activation gradients), where the only
remaining difference is FP32 accumulation order, and without rounding, which
bounds what bf16 costs."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from listctrecons_amd import recon

pytestmark = pytest.mark.gpu

C, T = 32, 27


def unpack(p):
    p = torch.as_tensor(p, dtype=torch.float64)
    off = 0

    def take(n):
        nonlocal off
        out = p[off:off + n]
        off += n
        return out

    W1 = take(T * C).reshape(3, 3, 3, C).permute(3, 0, 1, 2)[:, None]   # [co][1][kz][ky][kx]
    b1 = take(C)
    mids = []
    for _ in range(4):
        W = take(T * C * C).reshape(3, 3, 3, C, C).permute(3, 4, 0, 1, 2)  # [co][ci][k..]
        mids.append((W, take(C)))
    W6 = take(T * C).reshape(3, 3, 3, C).permute(3, 0, 1, 2)[None]     # [1][ci][k..]
    b6 = take(1)
    assert off == p.numel()
    return W1, b1, mids, W6, b6


def bf(t, on):
    return t.to(torch.bfloat16).to(torch.float64) if on else t


def reference(p, x, t, rounding):
    """Forward, MSE loss and parameter gradients, rounding where the kernels round."""
    W1, b1, mids, W6, b6 = unpack(p)
    x = torch.as_tensor(x, dtype=torch.float64)[:, None]
    t = torch.as_tensor(t, dtype=torch.float64)[:, None]
    W1r, W6r = bf(W1, rounding), bf(W6, rounding)
    hs = [bf(F.relu(F.conv3d(x, W1r, b1, padding=1)), rounding)]
    for W, b in mids:
        hs.append(bf(F.relu(F.conv3d(hs[-1], bf(W, rounding), b, padding=1)), rounding))
    y = x + F.conv3d(hs[-1], W6r, b6, padding=1)
    n = y.numel()
    loss = float(((y - t) ** 2).mean())
    dy = 2 * (y - t) / n
    grads = []
    gW6 = torch.nn.grad.conv3d_weight(hs[4], W6.shape, dy, padding=1)
    g = bf((F.conv_transpose3d(dy, W6r, padding=1)) * (hs[4] > 0), rounding)
    mid_grads = []
    for k in range(3, -1, -1):
        W, b = mids[k]
        gW = torch.nn.grad.conv3d_weight(hs[k], W.shape, g, padding=1)
        gb = g.sum(dim=(0, 2, 3, 4))
        mid_grads.append((gW, gb))
        g = bf(F.conv_transpose3d(g, bf(W, rounding), padding=1) * (hs[k] > 0), rounding)
    gW1 = torch.nn.grad.conv3d_weight(x, W1.shape, g, padding=1)
    gb1 = g.sum(dim=(0, 2, 3, 4))
    # repack in the kernels' order
    out = [gW1[:, 0].permute(1, 2, 3, 0).reshape(-1), gb1]
    for gW, gb in reversed(mid_grads):
        out += [gW.permute(2, 3, 4, 0, 1).reshape(-1), gb]
    out += [gW6[0].permute(1, 2, 3, 0).reshape(-1), dy.sum().reshape(1)]
    return y[:, 0].numpy(), loss, torch.cat(out).numpy()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def problem():
    rng = np.random.default_rng(0)
    x = rng.uniform(0.0, 1.0, size=(2, 12, 16, 32)).astype(np.float32)
    t = (x + 0.1 * rng.normal(size=x.shape)).astype(np.float32)
    return x, t


def test_parameter_packing_round_trip(ctx):
    net = recon.Cnn(ctx, seed=3)
    p = net.params()
    assert p.size == 27 * 32 + 32 + 4 * (27 * 32 * 32 + 32) + 27 * 32 + 1
    q = p * 1.5
    net.set_params(q)
    np.testing.assert_array_equal(net.params(), q)
    net.close()


def test_forward_and_gradients(ctx, problem):
    x, t = problem
    net = recon.Cnn(ctx, seed=1)
    p = net.params()
    # a non-trivial last layer so every gradient path carries signal
    p[-(27 * 32 + 1):] = np.random.default_rng(2).normal(0, 0.05, 27 * 32 + 1)
    net.set_params(p)
    dx, dt_ = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    y = net.forward(dx).cpu().numpy()
    loss = net.loss_grad(dx, dt_)
    g = net.grads()
    y_r, loss_r, g_r = reference(p, x, t, rounding=True)
    y_e, loss_e, g_e = reference(p, x, t, rounding=False)
    print(f"forward rel {rel(y, y_r):.2e} (exact {rel(y, y_e):.2e}); grad rel {rel(g, g_r):.2e} "
          f"(exact {rel(g, g_e):.2e}); loss {loss:.6g} ref {loss_r:.6g}", flush=True)
    # FP32 accumulation order flips a few bf16 roundings of activations; each
    # flip moves one activation by one bf16 ulp
    assert rel(y, y_r) < 5e-3
    assert loss == pytest.approx(loss_r, rel=1e-3)
    assert rel(g, g_r) < 2e-3
    assert rel(g, g_e) < 3e-2
    # every layer's block of the gradient agrees, not only the total
    offs = np.cumsum([0, 27 * 32, 32] + [27 * 32 * 32, 32] * 4 + [27 * 32, 1])
    for a, b in zip(offs[:-1], offs[1:]):
        assert rel(g[a:b], g_r[a:b]) < 5e-3, (a, b)
    net.close()


def test_adam_matches_the_update_rule(ctx, problem):
    x, t = problem
    net = recon.Cnn(ctx, seed=4)
    p0 = net.params().astype(np.float64)
    dx, dt_ = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    p = p0.copy()
    for step in range(1, 4):
        net.loss_grad(dx, dt_)
        g = net.grads().astype(np.float64)
        net.adam_step(1e-3)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p -= 1e-3 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
        np.testing.assert_allclose(net.params(), p, rtol=1e-5, atol=1e-6)
        p = net.params().astype(np.float64)  # follow the device's FP32 state
    net.close()


def test_gradients_are_bit_reproducible(ctx, problem):
    """Every reduction over voxels is a fixed-order sum of per-chunk
    partials: the same inputs give the same loss and gradients bit for bit,
    and two networks from one seed follow the same Adam trajectory."""
    x, t = problem
    dx, dt_ = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    nets = [recon.Cnn(ctx, seed=7) for _ in range(2)]
    l0 = nets[0].loss_grad(dx, dt_)
    g0 = nets[0].grads()
    l1 = nets[0].loss_grad(dx, dt_)
    assert l0 == l1
    np.testing.assert_array_equal(nets[0].grads(), g0)
    for net in nets:
        for _ in range(10):
            net.loss_grad(dx, dt_)
            net.adam_step(1e-3)
    np.testing.assert_array_equal(nets[0].params(), nets[1].params())
    for net in nets:
        net.close()


def test_training_reduces_the_loss(ctx):
    """Learn a local non-linear map (clip to [0.3, 0.7] of a smooth field):
    the loss must fall by 4x in 100 Adam steps.  The trajectory is
    deterministic (fixed-order reductions), so the ratio printed here is the
    ratio every run reaches."""
    rng = np.random.default_rng(5)
    noise = torch.from_numpy(rng.normal(size=(4, 1, 16, 16, 32)).astype(np.float32))
    x = F.avg_pool3d(noise, 5, stride=1, padding=2, count_include_pad=False)[:, 0]
    x = (x - x.mean()) / x.std() * 0.25 + 0.5
    t = x.clamp(0.3, 0.7)
    dx, dt_ = x.contiguous().cuda(), t.contiguous().cuda()
    net = recon.Cnn(ctx, seed=6)
    first = net.loss_grad(dx, dt_)
    for _ in range(100):
        net.adam_step(1e-3)
        last = net.loss_grad(dx, dt_)
    print(f"loss {first:.4g} -> {last:.4g} (ratio {first / last:.2f})", flush=True)
    assert last < 0.25 * first, f"loss fell {first / last:.2f}x in 100 steps"
    net.close()
