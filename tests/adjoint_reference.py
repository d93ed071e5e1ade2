"""Float64 reference for the sequential-lens spot RMS and its gradient:
the same model as rpx_lens_spot / rpx_lens_spot_grad (include/rpx.h),
written with torch tensor operations on the CPU and differentiated by torch
autograd (the sag's derivative in rho is itself taken by autograd, so no
hand-derived formula is shared with the kernels).  Each intersection is
solved by Newton's method on detached values and then refined by one tracked
Newton step, whose derivative is the implicit-function derivative of the
root.  Test infrastructure only."""
import numpy as np
import torch


def sag(P, rho):
    c, k, A4, A6, A8, A10 = P[0], P[1], P[2], P[3], P[4], P[5]
    w = torch.sqrt(1 - (1 + k) * c * c * rho)
    return c * rho / (1 + w) + A4 * rho ** 2 + A6 * rho ** 3 + A8 * rho ** 4 + A10 * rho ** 5


def sag_prime(P, rho):
    """dS/drho, differentiable in P and rho."""
    with torch.enable_grad():
        r = rho if rho.requires_grad else rho.detach().requires_grad_(True)
        (g,) = torch.autograd.grad(sag(P, r).sum(), r, create_graph=True)
    return g


def _step(P, zs, p, d, t):
    q = p + t[:, None] * d
    rho = q[:, 0] ** 2 + q[:, 1] ** 2
    G = sag(P, rho) - (q[:, 2] - zs)
    Gt = sag_prime(P, rho) * 2 * (q[:, 0] * d[:, 0] + q[:, 1] * d[:, 1]) - d[:, 2]
    return t - G / Gt


def spot_rms(params, lens, origins, directions, wl_idx, wavelengths):
    """params: (S, 6) float64 tensor; returns (spot RMS tensor, kept mask)."""
    table = torch.from_numpy(lens.index_table(wavelengths))
    p = torch.from_numpy(np.ascontiguousarray(origins, np.float64))
    d = torch.from_numpy(np.ascontiguousarray(directions, np.float64))
    d = d / torch.linalg.norm(d, dim=1, keepdim=True)
    nt = table[torch.from_numpy(np.asarray(wl_idx, np.int64))]
    keep = torch.ones(len(p), dtype=torch.bool)
    for s in range(lens.n_surfaces):
        P = params[s]
        zs = float(lens.z[s])
        t = ((zs - p[:, 2]) / d[:, 2]).detach()
        Pd = P.detach()
        for _ in range(60):
            t = _step(Pd, zs, p.detach(), d.detach(), t).detach()
        t = _step(P, zs, p, d, t)
        q = p + t[:, None] * d
        rho = q[:, 0] ** 2 + q[:, 1] ** 2
        keep = keep & (rho <= float(lens.apertures[s]) ** 2).detach()
        Sp = sag_prime(P, rho)
        m = torch.stack([-2 * q[:, 0] * Sp, -2 * q[:, 1] * Sp, torch.ones_like(Sp)], dim=1)
        N = m / torch.linalg.norm(m, dim=1, keepdim=True)
        sgn = torch.where((d * N).sum(1) >= 0, 1.0, -1.0).to(torch.float64).detach()
        N = N * sgn[:, None]
        mu = nt[:, lens.before[s]] / nt[:, lens.after[s]]
        cosv = (d * N).sum(1)
        r = 1 - mu * mu * (1 - cosv * cosv)
        keep = keep & (r >= 0).detach()
        r = torch.clamp(r, min=0.0)
        d = mu[:, None] * d + (torch.sqrt(r) - mu * cosv)[:, None] * N
        p = q
    t = (lens.detector_z - p[:, 2]) / d[:, 2]
    x = (p[:, 0] + t * d[:, 0])[keep]
    y = (p[:, 1] + t * d[:, 1])[keep]
    cx, cy = x.mean(), y.mean()
    return torch.sqrt(((x - cx) ** 2 + (y - cy) ** 2).mean()), keep


def spot_rms_grad(lens, origins, directions, wl_idx, wavelengths, params=None):
    P = torch.tensor(lens.params if params is None else params, dtype=torch.float64, requires_grad=True)
    sigma, keep = spot_rms(P, lens, origins, directions, wl_idx, wavelengths)
    (g,) = torch.autograd.grad(sigma, P)
    return float(sigma), g.numpy(), keep.numpy()


# ------------------------------------------------ distortions and objectives
MONO = [(a, d - a) for d in range(7) for a in range(d, -1, -1)]


def surface_height(P, C, ru, x, y):
    """Sag S(rho) plus the distortion sum_k C_k (x/ru)^a (y/ru)^b."""
    h = sag(P, x * x + y * y)
    if C is not None:
        xn, yn = x / ru, y / ru
        for k, (a, b) in enumerate(MONO):
            h = h + C[k] * xn ** a * yn ** b
    return h


def height_grad(P, C, ru, x, y):
    """(dh/dx, dh/dy), differentiable (autograd with create_graph)."""
    with torch.enable_grad():
        xr = x if x.requires_grad else x.detach().requires_grad_(True)
        yr = y if y.requires_grad else y.detach().requires_grad_(True)
        h = surface_height(P, C, ru, xr, yr).sum()
        gx, gy = torch.autograd.grad(h, (xr, yr), create_graph=True)
    return gx, gy


def _step2(P, C, ru, zs, p, d, t):
    q = p + t[:, None] * d
    G = surface_height(P, C, ru, q[:, 0], q[:, 1]) - (q[:, 2] - zs)
    gx, gy = height_grad(P, C, ru, q[:, 0], q[:, 1])
    Gt = gx * d[:, 0] + gy * d[:, 1] - d[:, 2]
    return t - G / Gt


def objective(params, ztable, lens, origins, directions, wl_idx, wavelengths, kind, acc0=None):
    """Spot RMS ("spot") or wavefront RMS ("wavefront") of the lens model of
    rpx_lens_eval; params (S, 6) and ztable (S, 28) are float64 tensors (the
    monomial table, zero rows without a distortion).  Returns (value, kept
    mask)."""
    table = torch.from_numpy(lens.index_table(wavelengths))
    p = torch.from_numpy(np.ascontiguousarray(origins, np.float64))
    d = torch.from_numpy(np.ascontiguousarray(directions, np.float64))
    d = d / torch.linalg.norm(d, dim=1, keepdim=True)
    wl = torch.from_numpy(np.asarray(wl_idx, np.int64))
    nt = table[wl]
    W = torch.zeros(len(p), dtype=torch.float64) if acc0 is None else torch.from_numpy(np.asarray(acc0, np.float64))
    keep = torch.ones(len(p), dtype=torch.bool)
    for s in range(lens.n_surfaces):
        P = params[s]
        ru = float(lens.zernike_radius[s])
        C = ztable[s] if ru > 0 else None
        zs = float(lens.z[s])
        t = ((zs - p[:, 2]) / d[:, 2]).detach()
        Pd = P.detach()
        Cd = C.detach() if C is not None else None
        for _ in range(60):
            t = _step2(Pd, Cd, ru, zs, p.detach(), d.detach(), t).detach()
        t = _step2(P, C, ru, zs, p, d, t)
        q = p + t[:, None] * d
        rho = q[:, 0] ** 2 + q[:, 1] ** 2
        keep = keep & (rho <= float(lens.apertures[s]) ** 2).detach()
        gx, gy = height_grad(P, C, ru, q[:, 0], q[:, 1])
        m = torch.stack([-gx, -gy, torch.ones_like(gx)], dim=1)
        N = m / torch.linalg.norm(m, dim=1, keepdim=True)
        sgn = torch.where((d * N).sum(1) >= 0, 1.0, -1.0).to(torch.float64).detach()
        N = N * sgn[:, None]
        mu = nt[:, lens.before[s]] / nt[:, lens.after[s]]
        cosv = (d * N).sum(1)
        r = 1 - mu * mu * (1 - cosv * cosv)
        keep = keep & (r >= 0).detach()
        r = torch.clamp(r, min=0.0)
        W = W + nt[:, lens.before[s]] * t
        d = mu[:, None] * d + (torch.sqrt(r) - mu * cosv)[:, None] * N
        p = q
    if kind == "spot":
        t = (lens.detector_z - p[:, 2]) / d[:, 2]
        x = (p[:, 0] + t * d[:, 0])[keep]
        y = (p[:, 1] + t * d[:, 1])[keep]
        cx, cy = x.mean(), y.mean()
        return torch.sqrt(((x - cx) ** 2 + (y - cy) ** 2).mean()), keep
    nL = nt[:, lens.after[lens.n_surfaces - 1]]
    W = W + nL * (-p[:, 0] * d[:, 0] - p[:, 1] * d[:, 1] + (lens.detector_z - p[:, 2]) * d[:, 2])
    Wk, wk = W[keep], wl[keep]
    dev = torch.zeros_like(Wk)
    for l in torch.unique(wk):
        sel = wk == l
        dev = torch.where(sel, Wk - Wk[sel].mean(), dev)
    return torch.sqrt((dev * dev).mean()), keep


def objective_grad(lens, origins, directions, wl_idx, wavelengths, kind, params=None, zcoefs=None, acc0=None):
    """(value, d/d params (S, 6), d/d zcoefs (flat, lens.zcoefs order), kept)."""
    P = torch.tensor(lens.params if params is None else params, dtype=torch.float64, requires_grad=True)
    zc = torch.tensor(lens.zcoefs if zcoefs is None else zcoefs, dtype=torch.float64, requires_grad=True)
    rows = []
    k = 0
    for s in range(lens.n_surfaces):
        T = torch.from_numpy(lens._zT[s])
        J = T.shape[0]
        rows.append(zc[k:k + J] @ T if J else torch.zeros(T.shape[1], dtype=torch.float64))
        k += J
    ztable = torch.stack(rows)
    val, keep = objective(P, ztable, lens, origins, directions, wl_idx, wavelengths, kind, acc0)
    gp, gz = torch.autograd.grad(val, (P, zc), allow_unused=True)
    gz = gz if gz is not None else torch.zeros_like(zc)
    return float(val), gp.numpy(), gz.numpy(), keep.numpy()
