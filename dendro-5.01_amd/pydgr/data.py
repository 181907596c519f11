"""Seeded synthetic initial data on padded blocks (torch, any device).

Every padded point of every block is filled from an analytic field, so the
ghost layers hold the same data a perfect unzip would give.
Variables are in VAR order (BSSN_GR/include/grDef.h:31).
"""
import math

import numpy as np
import torch

NUM_VARS = 24
ALPHA, CHI, K, GT0, BETA0, B0, SG0, SA0 = 0, 1, 2, 3, 6, 9, 12, 18


def coords(blocks, device="cpu"):
    """x, y, z of every padded point (each (dof,) float64) and its flat index,
    listed block by block in the order of `blocks`."""
    xs, ys, zs, idx = [], [], [], []
    # group consecutive runs of equal size so each run is one broadcast
    runs = []
    for b in blocks:
        key = tuple(b["sz"])
        if runs and runs[-1][0] == key:
            runs[-1][1].append(b)
        else:
            runs.append((key, [b]))
    # built where they are used: on a GPU only the block table crosses PCIe,
    # from pinned memory without waiting (a blocking copy on a stream that
    # shares a hardware queue with other work would wait for that work)
    f64 = dict(dtype=torch.float64, device=device)
    dev = torch.device(device)

    def up(t):
        if dev.type != "cuda":
            return t
        return t.pin_memory().to(dev, non_blocking=True)
    for (s0, s1, s2), bl in runs:
        k, j, i = torch.meshgrid(torch.arange(s2, **f64), torch.arange(s1, **f64), torch.arange(s0, **f64),
                                 indexing="ij")
        i, j, k = i.reshape(1, -1), j.reshape(1, -1), k.reshape(1, -1)
        pm = up(torch.tensor([b["pmin"] + b["pmax"] for b in bl], dtype=torch.float64))
        pmin, pmax = pm[:, 0:3], pm[:, 3:6]
        h = torch.stack([(pmax[:, a] - pmin[:, a]) / float(n - 1) for a, n in enumerate((s0, s1, s2))], dim=1)
        xs.append((pmin[:, 0:1] + i * h[:, 0:1]).reshape(-1))
        ys.append((pmin[:, 1:2] + j * h[:, 1:2]).reshape(-1))
        zs.append((pmin[:, 2:3] + k * h[:, 2:3]).reshape(-1))
        off = up(torch.tensor([b["offset"] for b in bl], dtype=torch.int64)).reshape(-1, 1)
        idx.append((off + torch.arange(s0 * s1 * s2, dtype=torch.int64, device=device).reshape(1, -1)).reshape(-1))
    return torch.cat(xs), torch.cat(ys), torch.cat(zs), torch.cat(idx)


def _zeros(dof, device, out):
    """`out` (24, dof) cleared, or a new zero array."""
    if out is None:
        return torch.zeros((NUM_VARS, dof), dtype=torch.float64, device=device)
    if tuple(out.shape) != (NUM_VARS, dof) or out.dtype != torch.float64:
        raise ValueError("out must be float64 of shape (24, %d)" % dof)
    return out.zero_()


def _flat(dof, device, out=None):
    U = _zeros(dof, device, out)
    U[ALPHA] = 1.0
    U[CHI] = 1.0
    U[SG0 + 0] = 1.0
    U[SG0 + 3] = 1.0
    U[SG0 + 5] = 1.0
    return U


def flat_bumps(blocks, dof, seed=0, amp=1e-3, width=0.1, nbumps=8, box=None, device="cpu", out=None, plan=None):
    """Flat space (alpha=chi=1, gt=delta, rest 0) plus, on every variable,
    the sum of `nbumps` Gaussians amp*exp(-|x-c|^2/width^2) with seeded
    centres uniform in the bounding box of the blocks.  Written into `out`
    (24, dof) when given; temporaries hold a few variables at a time.  With
    a plan of these blocks (on a GPU), one fused kernel (dgr_init_bumps)
    writes every padded point, in the same operation order."""
    if plan is not None:
        return _flat_bumps_plan(plan, blocks, dof, seed, amp, width, nbumps, box, device, out)
    x, y, z, idx = coords(blocks, device)
    if box is None:
        lo = [min(b["pmin"][d] for b in blocks) for d in range(3)]
        hi = [max(b["pmax"][d] for b in blocks) for d in range(3)]
    else:
        lo, hi = box
    rng = np.random.default_rng(seed)
    cen = torch.tensor(rng.uniform(lo, hi, size=(NUM_VARS, nbumps, 3)), dtype=torch.float64)
    if torch.device(device).type == "cuda":
        cen = cen.pin_memory().to(device, non_blocking=True)
    U = _flat(dof, device, out)
    iw2 = 1.0 / (width * width)
    # four variables at a time, bump after bump (the same sums in the same
    # order as one variable at a time)
    G = 4
    acc = torch.empty((G, x.numel()), dtype=torch.float64, device=device)
    for v in range(0, NUM_VARS, G):
        acc.zero_()
        for b in range(nbumps):
            c = cen[v:v + G, b, :]
            acc += torch.exp(-((x[None, :] - c[:, 0:1]) ** 2 + (y[None, :] - c[:, 1:2]) ** 2 +
                               (z[None, :] - c[:, 2:3]) ** 2) * iw2)
        U[v:v + G, idx] += amp * acc
    return U


def _flat_bumps_plan(plan, blocks, dof, seed, amp, width, nbumps, box, device, out):
    if box is None:
        lo = [min(b["pmin"][d] for b in blocks) for d in range(3)]
        hi = [max(b["pmax"][d] for b in blocks) for d in range(3)]
    else:
        lo, hi = box
    rng = np.random.default_rng(seed)
    cen = torch.tensor(rng.uniform(lo, hi, size=(NUM_VARS, nbumps, 3)), dtype=torch.float64)
    cen = cen.pin_memory().to(device, non_blocking=True)
    U = out if out is not None else torch.empty((NUM_VARS, dof), dtype=torch.float64, device=device)
    if tuple(U.shape) != (NUM_VARS, dof) or U.dtype != torch.float64:
        raise ValueError("out must be float64 of shape (24, %d)" % dof)
    stream = torch.cuda.current_stream(U.device)
    plan.init_bumps(U, cen, amp, width, stream=stream)
    # the centres stay alive until the kernel has read them
    cen.record_stream(stream)
    return U


def gauge_wave(blocks, dof, amp=0.01, length=None, noise=1e-6, seed=0, device="cpu"):
    """Gauge wave: alpha = 1 - A sin(2 pi x / L), physical metric
    diag(alpha^2, 1, 1), so chi = alpha^(-2/3), gt_xx = alpha^(4/3),
    gt_yy = gt_zz = alpha^(-2/3); K_xx = -A (2pi/L) cos(2pi x/L) with zero
    shift; traceless At and Gt^x = -d_x gt^xx from those; plus seeded white
    noise of size `noise` on all 24 variables."""
    x, y, z, idx = coords(blocks, device)
    if length is None:
        length = max(b["pmax"][0] for b in blocks) - min(b["pmin"][0] for b in blocks)
    k = 2.0 * math.pi / length
    a = 1.0 - amp * torch.sin(k * x)
    da = -amp * k * torch.cos(k * x)
    U = torch.zeros((NUM_VARS, dof), dtype=torch.float64, device=device)
    chi = a ** (-2.0 / 3.0)
    Kxx = da  # -(1/2alpha) d_t gamma_xx with gamma_xx = alpha^2 moving at unit speed
    Ktr = Kxx / (a * a)
    U[ALPHA, idx] = a
    U[CHI, idx] = chi
    U[K, idx] = Ktr
    U[GT0, idx] = (4.0 / 3.0) * a ** (-7.0 / 3.0) * da
    U[SG0 + 0, idx] = a ** (4.0 / 3.0)
    U[SG0 + 3, idx] = chi
    U[SG0 + 5, idx] = chi
    U[SA0 + 0, idx] = chi * (Kxx - a * a * Ktr / 3.0)
    U[SA0 + 3, idx] = chi * (-Ktr / 3.0)
    U[SA0 + 5, idx] = chi * (-Ktr / 3.0)
    if noise:
        g = torch.Generator(device="cpu").manual_seed(seed)
        U += noise * torch.randn(U.shape, generator=g, dtype=torch.float64).to(device)
    return U


def brill_lindquist(blocks, dof, m1=0.5, m2=0.5, sep=8.0, offset=(0.0137, 0.0211, 0.0173), device="cpu",
                    out=None):
    """Brill-Lindquist punctures on the x axis at +-sep/2 (shifted by
    `offset` so no grid point sits on a puncture): chi = psi^-4 with
    psi = 1 + m1/(2 r1) + m2/(2 r2), alpha = chi^(1/2), gt = delta, all
    else zero.  Written into `out` (24, dof) when given."""
    x, y, z, idx = coords(blocks, device)
    r1 = torch.sqrt((x - 0.5 * sep - offset[0]) ** 2 + (y - offset[1]) ** 2 + (z - offset[2]) ** 2)
    r2 = torch.sqrt((x + 0.5 * sep - offset[0]) ** 2 + (y - offset[1]) ** 2 + (z - offset[2]) ** 2)
    psi = 1.0 + m1 / (2.0 * r1) + m2 / (2.0 * r2)
    chi = psi ** -4
    U = _zeros(dof, device, out)
    U[CHI, idx] = chi
    U[ALPHA, idx] = torch.sqrt(chi)
    U[SG0 + 0, idx] = 1.0
    U[SG0 + 3, idx] = 1.0
    U[SG0 + 5, idx] = 1.0
    return U
