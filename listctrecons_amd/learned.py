"""Learned helical artifact correction (host orchestration of the lct_cnn_* kernels).

Training pairs are made by the method itself: a seeded random-ellipsoid
phantom is projected analytically on the helical scan at pitch 1.5
(lct_project_ellipsoids + lct_poisson_noise), reconstructed by plain FDK
(lct_reconstruct), and paired with the voxelised phantom in the
reconstruction's scale (RECON_SCALE x mu).  The network y = x + net(x) is
trained on random 64^3 patches with Adam.  A corrected volume is accepted only
if its reprojection residual against the measured projections is no worse
than the FDK volume's; otherwise FDK is returned.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import _lib, recon, scan

NORM = 0.05  # volumes are divided by this before the network (2 x 0.025 /mm)


def learned_workload(readings: int = 6000, nx: int = 256, **kw) -> scan.Workload:
    """Plain-FDK geometry (16 x 368, 576 views/turn) at pitch 1.5; Hann window."""
    return scan.workload("helical_16x368_p15_256x128", readings=readings, nx=nx, **kw)


def phantom(seed: int, z_range, radius: float = 85.0) -> np.ndarray:
    """20-60 random ellipsoids (semi-axes 4-45 mm, attenuation U(0, 0.025)/mm)
    inside a `radius` mm cylinder over z_range."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 61))
    out = []
    for _ in range(n):
        axes = rng.uniform(4.0, 45.0, 3)
        r = max(0.0, radius - axes.max()) * math.sqrt(rng.uniform())
        a = rng.uniform(0, 2 * math.pi)
        zc = rng.uniform(*z_range)
        out.append(scan.make_ellipsoid([r * math.cos(a), r * math.sin(a), zc], axes,
                                       scan._random_rotation(rng), rng.uniform(0.0, 0.025)))
    return np.stack(out)


@dataclasses.dataclass
class Case:
    fdk: object        # torch [slices][ny][nx], reference scale
    target: object     # torch, RECON_SCALE x phantom density
    fan: object        # measured projections (device)
    plan: object
    angles: np.ndarray
    couch: np.ndarray
    z_ref: list        # table position of slice 0 of each block
    ell: np.ndarray


def make_case(ctx: recon.Context, w: scan.Workload, seed: int, preset: str = "mixed_f16") -> Case:
    import torch
    a, c = w.angles_couch()
    ell = phantom(seed, (float(c[0]), float(c[-1])))
    g = recon.scan_geometry(w)
    fan = torch.empty((w.readings, w.nt, w.ns), dtype=torch.float32, device="cuda")
    ctx.project_ellipsoids(g, a, c, ell, fan)
    ctx.poisson_noise(fan, w.i0, 10_000 + seed)
    pl = recon.plan_helical(w.prm(), a, c)
    vol = ctx.reconstruct(pl, recon.preset(preset, w.window), a, c, fan)
    target = torch.empty_like(vol)
    z_ref = []
    ig = pl.ig
    for b in range(pl.n_blocks):
        ref_view = recon.block_info(pl, c, b)[4]
        z0 = float(c[ref_view])
        z_ref.append(z0)
        ctx.voxelize(ig, z0, 32, ell, target[32 * b:32 * (b + 1)], supersample=2)
    target.mul_(scan.RECON_SCALE)
    ctx.synchronize()
    return Case(vol, target, fan, pl, a, c, z_ref, ell)


def random_patches(cases, batch: int, size: int, rng):
    import torch
    xs, ts = [], []
    for _ in range(batch):
        cs = cases[int(rng.integers(len(cases)))]
        D, H, W = cs.fdk.shape
        z = int(rng.integers(0, D - size + 1))
        y = int(rng.integers(0, H - size + 1))
        x = int(rng.integers(0, W - size + 1))
        xs.append(cs.fdk[z:z + size, y:y + size, x:x + size])
        ts.append(cs.target[z:z + size, y:y + size, x:x + size])
    return (torch.stack(xs).contiguous() / NORM, torch.stack(ts).contiguous() / NORM)


def train(ctx, net: recon.Cnn, cases, steps: int, batch: int = 8, size: int = 64,
          lr: float = 1e-3, seed: int = 0, log=None):
    rng = np.random.default_rng(seed)
    losses = []
    t0 = time.perf_counter()
    for s in range(steps):
        x, t = random_patches(cases, batch, size, rng)
        losses.append(net.loss_grad(x, t))
        net.adam_step(lr)
        if log and (s % 50 == 0 or s == steps - 1):
            log(f"step {s}: loss {losses[-1]:.5g} ({time.perf_counter() - t0:.1f} s)")
    ctx.synchronize()
    return losses


def correct(net: recon.Cnn, vol):
    """Apply y = x + net(x) to a whole volume [D][H][W] (reference scale)."""
    x = (vol / NORM)[None].contiguous()
    y = net.forward(x)
    return y[0] * NORM


def fov_mask(case: Case, radius: float = 90.0):
    import torch
    ig = case.plan.ig
    yy, xx = np.mgrid[0:ig.ny, 0:ig.nx]
    m = np.hypot((xx - ig.wx) * ig.dx, (yy - ig.wy) * ig.dy) < radius
    return torch.from_numpy(m).cuda()


def rmse_to_phantom(vol, case: Case) -> float:
    m = fov_mask(case)
    d = (vol - case.target)[:, m]
    return float(d.pow(2).mean().sqrt()) / scan.RECON_SCALE


def reprojection_residual(ctx, vol, case: Case, w: scan.Workload, margin: float = 10.0) -> float:
    """|| A(vol / RECON_SCALE) - p || / || p || over the views whose rays stay
    inside the reconstructed slab (source z at least `margin` mm inside it)."""
    import torch
    c = case.couch
    ig = case.plan.ig
    z_lo = case.z_ref[0]
    z_hi = case.z_ref[-1] + 31 * ig.dz
    views = np.nonzero((c >= z_lo + margin) & (c <= z_hi - margin))[0]
    if views.size == 0:
        raise ValueError("slab too thin for a reprojection test")
    v0, nv = int(views[0]), int(views[-1] - views[0] + 1)
    g = recon.scan_geometry(w)
    out = torch.empty((nv, w.nt, w.ns), dtype=torch.float32, device="cuda")
    # block slices are contiguous in z when the blocks abut
    ctx.forward_project(g, case.angles, c, (vol / scan.RECON_SCALE).contiguous(), ig, z_lo, out,
                        view0=v0, n_views=nv, step=0.4)
    ctx.synchronize()
    p = case.fan[v0:v0 + nv]
    return float((out - p).norm() / p.norm())


def gated(ctx, net, case: Case, w: scan.Workload):
    """Corrected volume if its reprojection residual is no worse than FDK's."""
    corr = correct(net, case.fdk)
    r_fdk = reprojection_residual(ctx, case.fdk, case, w)
    r_cor = reprojection_residual(ctx, corr, case, w)
    return (corr if r_cor <= r_fdk else case.fdk), r_fdk, r_cor


def run(train_phantoms: int = 2000, pool: int = 96, steps: int = 2000, refresh_every: int = 10,
        held_out: int = 2, seed: int = 0, log=print) -> dict:
    """The pitch-1.5 learned-correction experiment.

    Training: `train_phantoms` distinct seeded phantoms (seeds 1000...) are
    streamed through a pool of `pool` FDK/phantom pairs at 256 x 256 x 128,
    one pair replaced every `refresh_every` steps after the pool is full;
    batch 8 of 64^3 patches, Adam lr 1e-3.
    Evaluation: `held_out` unseen phantoms (seeds 9000...) reconstructed at
    512 x 512 x 256 (0.8 mm voxels, the training voxel size: the 512^2 grid
    spans the 208 mm scan field of view with margin), RMSE to the phantom in
    the field of view for FDK and for the gated correction, and both
    reprojection residuals."""
    import torch

    ctx = recon.Context(0)
    w = learned_workload()
    w_eval = learned_workload(readings=9900, nx=512)
    t0 = time.perf_counter()
    cases = [make_case(ctx, w, 1000 + i) for i in range(min(pool, train_phantoms))]
    gen_s = time.perf_counter() - t0
    next_seed = 1000 + len(cases)
    log(f"pool of {len(cases)} pairs in {gen_s:.1f} s")
    net = recon.Cnn(ctx, seed=seed)
    rng = np.random.default_rng(seed)
    losses = []
    t_train = time.perf_counter()
    for s in range(steps):
        if (s + 1) % refresh_every == 0 and next_seed < 1000 + train_phantoms:
            cases[int(rng.integers(len(cases)))] = make_case(ctx, w, next_seed)
            next_seed += 1
        x, t = random_patches(cases, 8, 64, rng)
        losses.append(net.loss_grad(x, t))
        net.adam_step(1e-3)
        if s % 100 == 0 or s == steps - 1:
            log(f"step {s}: loss {losses[-1]:.5g}  phantoms seen {next_seed - 1000}  "
                f"({time.perf_counter() - t_train:.0f} s)")
    ctx.synchronize()
    train_s = time.perf_counter() - t_train
    del cases
    torch.cuda.empty_cache()

    evals = []
    for k in range(held_out):
        c = make_case(ctx, w_eval, 9000 + k)
        torch.cuda.synchronize()
        ti = time.perf_counter()
        corr = correct(net, c.fdk)
        torch.cuda.synchronize()
        infer_s = time.perf_counter() - ti
        r_fdk = reprojection_residual(ctx, c.fdk, c, w_eval)
        r_cor = reprojection_residual(ctx, corr, c, w_eval)
        e_fdk = rmse_to_phantom(c.fdk, c)
        e_cor = rmse_to_phantom(corr, c)
        accepted = r_cor <= r_fdk
        evals.append({"seed": 9000 + k, "rmse_fdk_per_mm": e_fdk, "rmse_corrected_per_mm": e_cor,
                      "rmse_gain": 1 - e_cor / e_fdk,
                      "rmse_fdk_hu": e_fdk / scan.MU_WATER * 1000,
                      "rmse_corrected_hu": e_cor / scan.MU_WATER * 1000,
                      "reprojection_residual_fdk": r_fdk, "reprojection_residual_corrected": r_cor,
                      "gate_accepts_correction": accepted, "inference_s": infer_s})
        log(f"held-out {9000 + k}: rmse {e_fdk:.3e} -> {e_cor:.3e} ({1 - e_cor / e_fdk:.1%}), "
            f"reprojection {r_fdk:.4f} -> {r_cor:.4f}, accepted={accepted}, "
            f"inference {infer_s * 1e3:.0f} ms")
        del c, corr
        torch.cuda.empty_cache()
    net.close()
    ctx.close()
    return {"workload": "pitch 1.5, 16x368 detector, 576 views/turn",
            "train_volume": "256x256x128 @ 0.8 mm", "eval_volume": "512x512x256 @ 0.8 mm",
            "train_phantoms": next_seed - 1000, "pool": pool, "steps": steps, "batch": 8,
            "patch": 64, "lr": 1e-3, "pool_generation_s": gen_s, "train_s": train_s,
            "loss_first": float(np.mean(losses[:10])), "loss_last": float(np.mean(losses[-50:])),
            "held_out": evals}


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-phantoms", type=int, default=2000)
    ap.add_argument("--pool", type=int, default=96)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--refresh-every", type=int, default=1)
    ap.add_argument("--held-out", type=int, default=2)
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    res = run(args.train_phantoms, args.pool, args.steps, args.refresh_every, args.held_out)
    text = json.dumps(res, indent=1)
    print(text, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
