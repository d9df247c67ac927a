"""Float64 CPU restatement of the coupled step in torch, for fully periodic
2D (D2Q9) domains: the tape whose autograd gradients check the hand-written
adjoint.  Follows the reference form (two-pass WMRT collision with the
filter M^-1 S M, NsAllen.cu:369-763; streaming and the lagged phi ghost ring,
NsAllen.cu:881-892; initial state, NsAllen.cu:137-202), written with the
oracle's lattice tables.  Test infrastructure only.
"""
import numpy as np
import torch

from tests import oracle_lib


class TorchStep:
    def __init__(self, cfg, normal_floor=0.0):
        assert cfg.dim == 2
        self.nfloor = normal_floor
        lat = oracle_lib.lattice(2)
        self.e = torch.tensor(lat["e"][:, :2], dtype=torch.float64)
        self.ei = lat["e"][:, :2].astype(int)
        self.w = torch.tensor(lat["w"], dtype=torch.float64)
        self.M = torch.tensor(lat["M"], dtype=torch.float64)
        self.Minv = torch.tensor(lat["Minv"], dtype=torch.float64)
        self.stress = [7, 8]
        self.c = cfg
        self.drho = cfg.rho_h - cfg.rho_l

    # interior field x (nx, ny) with the ghost ring of g (periodic wrap of g)
    @staticmethod
    def pad(x, g):
        gp = torch.cat([g[-1:], g, g[:1]], 0)
        gp = torch.cat([gp[:, -1:], gp, gp[:, :1]], 1)
        P = gp.clone()
        P[1:-1, 1:-1] = x
        return P

    def stencil(self, P):
        c0 = P[1:-1, 1:-1]
        xp, xm = P[2:, 1:-1], P[:-2, 1:-1]
        yp, ym = P[1:-1, 2:], P[1:-1, :-2]
        pp, pm, mp, mm = P[2:, 2:], P[2:, :-2], P[:-2, 2:], P[:-2, :-2]
        dx = (xp - xm) / 3 + ((pm + pp) - (mm + mp)) / 12
        dy = (yp - ym) / 3 + ((mp + pp) - (mm + pm)) / 12
        lap = (((mm + mp) + pm) + pp + 4 * (((xm + ym) + yp) + xp) - 20 * c0) / 6
        return c0, torch.stack([dx, dy]), lap

    def gamma(self, u):
        eu = torch.einsum("qa,axy->qxy", self.e, u)
        uu = (u * u).sum(0)
        return self.w[:, None, None] * (3 * eu + 4.5 * eu * eu - 1.5 * uu)

    def filt(self, v, s):
        m = torch.einsum("rq,qxy->rxy", self.M, v)
        m = torch.stack([m[r] * s if r in self.stress else m[r] for r in range(9)])
        return torch.einsum("qr,rxy->qxy", self.Minv, m)

    def visc(self, gf, tau, dc):
        Pi = torch.einsum("qa,qb,qxy->abxy", self.e, self.e, gf)
        return -tau * torch.einsum("abxy,bxy->axy", Pi, dc) * self.drho

    def scalars(self, c0, lap, beta, kappa):
        c = self.c
        rho = c.rho_l + c0 * self.drho
        tau = torch.where(c0 < 0, torch.full_like(c0, c.tau_l),
                          torch.where(c0 > 1, torch.full_like(c0, c.tau_h), c.tau_l + c0 * (c.tau_h - c.tau_l)))
        s = 1 / (tau + 0.5)
        mu = 4 * beta * c0 * (c0 - 1) * (c0 - 0.5) - kappa * lap
        return rho, tau, s, mu

    def normal(self, dc, c0):
        """n = dc / (|dc| + 1e-32); with a floor c > 0 its derivative is scaled
        by |dc| / max(|dc|, c 4 phi (1 - phi) / W) (same value)."""
        if self.nfloor > 0:
            mag = torch.sqrt((dc * dc).sum(0))
            fl = (self.nfloor * 4 * c0 * (1 - c0) / self.c.interface_width).abs()
            sc = torch.where(mag > fl, torch.ones_like(mag), torch.where(fl > 0, mag / fl, torch.zeros_like(mag)))
            dc = dc.detach() + (dc - dc.detach()) * sc.detach()
        t = torch.sqrt((dc * dc).sum(0)) + 1e-32
        return dc / t

    def collide(self, fp, gp, p, u, c0, dc, lap, beta, kappa, mob):
        rho, tau, s, mu = self.scalars(c0, lap, beta, kappa)
        nr = self.normal(dc, c0)
        wc = 1 / (0.5 + 3 * mob)
        A = 3 * mob * (1 - 4 * (c0 - 0.5) ** 2) / self.c.interface_width
        en = torch.einsum("qa,axy->qxy", self.e, nr)
        Gm = self.gamma(u)
        heq = c0 * (Gm + self.w[:, None, None]) + self.w[:, None, None] * (A * en)
        g_new = gp * (1 - wc) + heq * wc
        gneq = fp - (p * self.w[:, None, None] + Gm)
        gf = self.filt(gneq, s)
        F = mu * dc - p * self.drho / 3 * dc + self.visc(gf, tau, dc)
        hlp = 3 * self.w[:, None, None] * torch.einsum("qa,axy->qxy", self.e, F) / rho
        gneq2 = fp - (p * self.w[:, None, None] + Gm - 0.5 * hlp)
        f_new = fp - self.filt(gneq2, s) + hlp
        return f_new, g_new

    def init(self, phi0, sigma, mob):
        c = self.c
        W = c.interface_width
        beta, kappa = 8 * sigma / W, 1.5 * sigma * W
        P = self.pad(phi0, phi0)
        c0, dc, lap = self.stencil(P)
        rho = c.rho_l + c0 * self.drho
        k = 1.0 if c.case_type == 1 else (-1.0 if c.case_type == 0 else 0.0)
        p = k * c0 * sigma / (c.reference_length / 2) / (rho / 3)
        nr = self.normal(dc, c0)
        en = torch.einsum("qa,axy->qxy", self.e, nr)
        w = self.w[:, None, None]
        f = p * w
        g = c0 * w - 0.5 * w * ((1 - 4 * (c0 - 0.5) ** 2) / W * en)
        u = torch.zeros_like(dc)
        F, G = self.collide(f, g, p, u, c0, dc, lap, beta, kappa, mob)
        return F, G, u, phi0

    def step(self, state, sigma, mob):
        F, G, U, Phi = state
        W = self.c.interface_width
        beta, kappa = 8 * sigma / W, 1.5 * sigma * W
        fp = torch.stack([torch.roll(F[q], shifts=(int(self.ei[q, 0]), int(self.ei[q, 1])), dims=(0, 1))
                          for q in range(9)])
        gp = torch.stack([torch.roll(G[q], shifts=(int(self.ei[q, 0]), int(self.ei[q, 1])), dims=(0, 1))
                          for q in range(9)])
        phi = gp.sum(0)
        c0, dc, lap = self.stencil(self.pad(phi, Phi))
        p = fp.sum(0)
        rho, tau, s, mu = self.scalars(c0, lap, beta, kappa)
        gneq = fp - (p * self.w[:, None, None] + self.gamma(U))
        gf = self.filt(gneq, s)
        F_ = mu * dc - p * self.drho / 3 * dc + self.visc(gf, tau, dc)
        u = torch.einsum("qa,qxy->axy", self.e, fp) + 0.5 * F_ / rho
        f_new, g_new = self.collide(fp, gp, p, u, c0, dc, lap, beta, kappa, mob)
        return f_new, g_new, u, phi

    def run(self, phi0, sigma, mob, nsteps):
        st = self.init(phi0, sigma, mob)
        for _ in range(nsteps):
            st = self.step(st, sigma, mob)
        return st


def interior_phi0(cfg):
    """The case's initial phi on the interior nodes, [i, j]."""
    from tests.oracle_lib import Oracle
    return np.ascontiguousarray(Oracle(cfg).init().interior("phi")[..., 0, 0])


def wrapped_grid(phi0):
    """phi on the 2D grid (ext_x, ext_y, 1, 1) from interior values, ghosts by periodic wrap."""
    p = np.pad(phi0, 1, mode="wrap")
    return p[:, :, None, None]


def perturbed(phi0, seed=0, amp=1e-3):
    """phi0 + amp N(0,1) noise smoothed by one 5-point Laplacian pass,
    confined to the interface band by 4 phi0 (1 - phi0)."""
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal(phi0.shape)
    lap = sum(np.roll(xi, s, a) for a in range(phi0.ndim) for s in (1, -1)) - 2 * phi0.ndim * xi
    return phi0 + amp * (xi + 0.25 * lap) * 4 * phi0 * (1 - phi0)
