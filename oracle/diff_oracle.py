"""FP64 restatement of the differentiable frozen-topology MC-Glauber batch
(test infrastructure: used only by tests/ as the checker of the HIP
adjoint).  Independent numerics from the device code: the Woods-Saxon
radial quantile uses a 256-node Gauss-Legendre rule (pinned against scipy
quad + brentq), the Gamma quantile is scipy's gammaincinv, and the adjoint
is written in vectorised numpy.

Model (theta = w, R, a, alpha, k), for one event of a frozen topology:
  r_i = F^{-1}(u_i; R, a), F(r) = N(r)/N(R + 2.5), N(r) = int_0^r t^2/(1+e^{(t-R)/a}) dt
      (the radial law of Nucleus.cpp:578-621)
  (x, y)_i = r_i (sin th_i cos ph_i, sin th_i sin ph_i) - CM + (+-b/2, 0)   (Nucleus.cpp:296-309)
  sources: participants (wound order, projectile then target), then pair
      midpoints (MCnucl.cpp:326-352); weights (1-alpha)/2 f and alpha f with
      f = G^{-1}(v; s)/s, s = (1-alpha+1e-8)k/2 or (alpha+1e-8)k (MCnucl.cpp:1271-1301)
  rho = sum_k W_k exp(-d^2/2w^2)/(2 pi w^2) over each source's integer window
      frozen at theta0 (participant box, +-5w window; MCnucl.cpp:450-475, 822-866)
  eps_n = -sum_mask rho z^n r^(2-n) / sum_mask rho r^2 about the centre of
      mass (MakeDensity.cpp:2273-2423), mask frozen at theta0
  outputs |eps_2|, |eps_3|, dS/dy = finalFactor dx dy sum rho
"""
import numpy as np
from scipy import special

import oracle

KQ = 2.0 ** -24
_X, _W = np.polynomial.legendre.leggauss(256)
_X = 0.5 * (_X + 1.0)
_W = 0.5 * _W


def _ws_ints(r, R, a):
    t = r[:, None] * _X[None, :]
    wt = r[:, None] * _W[None, :]
    rho = 1.0 / (1.0 + np.exp((t - R) / a))
    dR = rho * (1.0 - rho) / a
    t2w = t * t * wt
    return (t2w * rho).sum(1), (t2w * dR).sum(1), (t2w * dR * (t - R) / a).sum(1)


def ws_finv(u, R, a):
    """r = F^{-1}(u; R, a) and dr/dR, dr/da (implicit differentiation)."""
    u = np.atleast_1d(np.asarray(u, np.float64))
    rmax = R + 2.5
    Z, ZR, Za = (v[0] for v in _ws_ints(np.array([rmax]), R, a))
    ZR = ZR + rmax * rmax / (1.0 + np.exp(2.5 / a))
    lo = np.zeros_like(u)
    hi = np.full_like(u, rmax)
    r = rmax * np.cbrt(u)
    for _ in range(200):
        n = _ws_ints(r, R, a)[0]
        f = n - u * Z
        lo = np.where(f < 0, r, lo)
        hi = np.where(f < 0, hi, r)
        rho = 1.0 / (1.0 + np.exp((r - R) / a))
        rn = r - f / (r * r * rho)
        bad = ~((rn > lo) & (rn < hi))
        rn = np.where(bad, 0.5 * (lo + hi), rn)
        done = np.abs(rn - r) <= 1e-15 * rmax
        r = rn
        if done.all():
            break
    n, nR, na = _ws_ints(r, R, a)
    rho = 1.0 / (1.0 + np.exp((r - R) / a))
    Fr = r * r * rho / Z
    FR = nR / Z - n * ZR / Z ** 2
    Fa = na / Z - n * Za / Z ** 2
    return r, -FR / Fr, -Fa / Fr


def gamma_dP_ds(s, x):
    """dP(s, x)/ds of the regularised lower incomplete gamma, by the series
    P = e^-x x^s/Gamma(s+1) sum_n x^n/((s+1)..(s+n))."""
    s = np.asarray(s, np.float64)
    x = np.asarray(x, np.float64)
    lx = np.log(x)
    pref = np.exp(-x + s * lx - special.gammaln(s + 1.0))
    t = np.ones_like(x)
    tot = np.ones_like(x)
    H = np.zeros_like(x)
    dsum = np.zeros_like(x)
    n = 1
    while True:
        sn = s + n
        t = t * x / sn
        H = H + 1.0 / sn
        tot = tot + t
        dsum = dsum + t * H
        if n > x.max() and (t < 1e-18 * tot).all():
            break
        n += 1
    P = pref * tot
    return P * (lx - special.digamma(s + 1.0)) - pref * dsum


def gamma_quantile(v, s):
    """Unit-scale Gamma quantile x(v; s) and dx/ds."""
    x = special.gammaincinv(s, v)
    pdf = np.exp(-x + (s - 1.0) * np.log(x) - special.gammaln(s))
    return x, -gamma_dP_ds(s, x) / pdf


def shapes(alpha, k):
    """Gamma shapes of participants and pairs (MCnucl.cpp:1271-1301)."""
    return (1.0 - alpha + 1e-8) * 0.5 * k, (alpha + 1e-8) * k


class DiffEvent:
    """One event of a frozen topology.  uvw_*: [A][3] (u, cos theta, phi) in
    collision order; xy0_*: [A][2] positions at theta0 as the topology was
    built (quantised); wound_*: participant rank or -1; pairs: (i, j) in pair
    order; gamma_v: Gamma uniforms per source (participants in wound order,
    projectile then target, then pairs); w0: width at theta0."""

    def __init__(self, g, uvw_a, uvw_b, b, xy0_a, xy0_b, wound_a, wound_b, pairs, gamma_v, w0, ff, fluct=True):
        self.g, self.ff, self.fluct = g, ff, fluct
        self.uvw = [np.asarray(uvw_a, np.float64), np.asarray(uvw_b, np.float64)]
        self.b = b
        self.wound = [np.asarray(wound_a), np.asarray(wound_b)]
        self.part = [np.argsort(np.where(w >= 0, w, 1 << 30))[: (w >= 0).sum()] for w in self.wound]
        self.pairs = np.asarray(pairs, int).reshape(-1, 2)
        self.gamma_v = np.asarray(gamma_v, np.float64)
        xy0 = [np.asarray(xy0_a, np.float64), np.asarray(xy0_b, np.float64)]
        # frozen windows: participant boxes, +-5w0 around pair midpoints
        self.win = []
        for nu in range(2):
            for i in self.part[nu]:
                bx = oracle.particle_box(xy0[nu][i, 0], xy0[nu][i, 1], w0)
                self.win.append(self._clamp(int((bx[0] - g.xmin) / g.dx), int((bx[1] - g.xmin) / g.dx),
                                            int((bx[2] - g.ymin) / g.dy), int((bx[3] - g.ymin) / g.dy)))
        for i, j in self.pairs:
            xm = 0.5 * (xy0[0][i, 0] + xy0[1][j, 0])
            ym = 0.5 * (xy0[0][i, 1] + xy0[1][j, 1])
            self.win.append(self._clamp(int((xm - 5 * w0 - g.xmin) / g.dx), int((xm + 5 * w0 - g.xmin) / g.dx),
                                        int((ym - 5 * w0 - g.ymin) / g.dy), int((ym + 5 * w0 - g.ymin) / g.dy)))
        boxes = [oracle.particle_box(xy0[nu][i, 0], xy0[nu][i, 1], w0) for nu in range(2) for i in self.part[nu]]
        self.mask = oracle.hotspot_mask(g, np.array(boxes).reshape(-1, 4), len(self.pairs))
        self.X = (g.xmin + g.dx * np.arange(g.nx))[:, None] * np.ones((1, g.ny))
        self.Y = np.ones((g.nx, 1)) * (g.ymin + g.dy * np.arange(g.ny))[None, :]

    def _clamp(self, x0, x1, y0, y1):
        g = self.g
        return max(0, x0), min(g.nx, x1), max(0, y0), min(g.ny, y1)

    def positions(self, R, a):
        """per nucleus: xy [A][2] and d(xy)/dR, d(xy)/da [A][2]."""
        out = []
        for nu in range(2):
            u, c, ph = self.uvw[nu].T
            r, rR, ra = ws_finv(u, R, a)
            st = np.sqrt(np.maximum(0.0, 1.0 - c * c))
            n = np.stack([st * np.cos(ph), st * np.sin(ph)], 1)
            p, pR, pa = r[:, None] * n, rR[:, None] * n, ra[:, None] * n
            xy = p - p.mean(0)
            xy[:, 0] += (0.5 if nu == 0 else -0.5) * self.b
            out.append((xy, pR - pR.mean(0), pa - pa.mean(0)))
        return out

    def sources(self, th):
        w, R, a, alpha, k = th
        pos = self.positions(R, a)
        xs, ps = [], []
        for nu in range(2):
            for i in self.part[nu]:
                xs.append(pos[nu][0][i])
                ps.append(0)
        for i, j in self.pairs:
            xs.append(0.5 * (pos[0][0][i] + pos[1][0][j]))
            ps.append(1)
        xs = np.array(xs).reshape(-1, 2)
        ps = np.array(ps)
        sp, sc = shapes(alpha, k)
        sh = np.where(ps == 1, sc, sp)
        if self.fluct:
            q, dq = gamma_quantile(self.gamma_v, sh)
            f, fs = q / sh, dq / sh - q / sh ** 2
        else:
            f, fs = np.ones(len(ps)), np.zeros(len(ps))
        pref = np.where(ps == 1, alpha, 0.5 * (1.0 - alpha))
        return pos, xs, ps, f, fs, pref

    def density(self, th):
        w = th[0]
        pos, xs, ps, f, fs, pref = self.sources(th)
        rho = np.zeros((self.g.nx, self.g.ny))
        norm = 1.0 / (2 * np.pi * w * w)
        if th[3] <= 1e-8:
            keep = ps == 0
        else:
            keep = np.ones(len(ps), bool)
        for kk, (x0, x1, y0, y1) in enumerate(self.win):
            if not keep[kk] or x0 >= x1 or y0 >= y1:
                continue
            d2 = (self.X[x0:x1, y0:y1] - xs[kk, 0]) ** 2 + (self.Y[x0:x1, y0:y1] - xs[kk, 1]) ** 2
            rho[x0:x1, y0:y1] += pref[kk] * f[kk] * norm * np.exp(-d2 / (2 * w * w))
        return rho

    def sums(self, rho):
        S = rho.sum()
        xc = (rho * self.X).sum() / S
        yc = (rho * self.Y).sum() / S
        x, y = self.X - xc, self.Y - yc
        m = self.mask & (rho != 0)
        r2 = x * x + y * y
        m &= r2 > 0
        z = x + 1j * y
        rr = np.sqrt(np.where(m, r2, 1.0))
        A2 = (rho * z ** 2)[m].sum()
        A3 = (rho * z ** 3 / rr)[m].sum()
        N = (rho * r2)[m].sum()
        return dict(S=S, xc=xc, yc=yc, x=x, y=y, z=z, r2=r2, rr=rr, m=m, A2=A2, A3=A3, N=N)

    def forward(self, th):
        rho = self.density(th)
        q = self.sums(rho)
        N = q["N"] + 1e-15
        return np.array([abs(q["A2"]) / N, abs(q["A3"]) / N, q["S"] * self.ff * self.g.dx * self.g.dy])

    def adjoint(self, th, cot):
        """dL/dtheta for L = cot . forward(theta)."""
        w, R, a, alpha, k = th
        rho = self.density(th)
        q = self.sums(rho)
        S, N, m, x, y, z, rr, r2 = q["S"], q["N"] + 1e-15, q["m"], q["x"], q["y"], q["z"], q["rr"], q["r2"]
        G = np.full(rho.shape, cot[2] * self.ff * self.g.dx * self.g.dy)
        Mx, My = (rho * x)[m].sum(), (rho * y)[m].sum()
        dNx, dNy = -2 * Mx, -2 * My
        for n, A, lam in ((2, q["A2"], cot[0]), (3, q["A3"], cot[1])):
            if lam == 0 or abs(A) == 0:
                continue
            if n == 2:
                dAx, dAy = -2 * (Mx + 1j * My), -2j * (Mx + 1j * My)
                gz = z ** 2
            else:
                dgx = 3 * z ** 2 / rr - z ** 3 * x / rr ** 3
                dgy = 3j * z ** 2 / rr - z ** 3 * y / rr ** 3
                dAx, dAy = -(rho * dgx)[m].sum(), -(rho * dgy)[m].sum()
                gz = z ** 3 / rr
            cx = (A.conjugate() * dAx).real / (abs(A) * N) - abs(A) * dNx / N ** 2
            cy = (A.conjugate() * dAy).real / (abs(A) * N) - abs(A) * dNy / N ** 2
            v = (cx * x + cy * y) / S
            v = v + np.where(m, (A.conjugate() * gz).real / (abs(A) * N) - abs(A) * r2 / N ** 2, 0.0)
            G += lam * v
        pos, xs, ps, f, fs, pref = self.sources(th)
        norm = 1.0 / (2 * np.pi * w * w)
        K = len(ps)
        gW, gX, gY, gw = np.zeros(K), np.zeros(K), np.zeros(K), 0.0
        for kk, (x0, x1, y0, y1) in enumerate(self.win):
            if x0 >= x1 or y0 >= y1 or (alpha <= 1e-8 and ps[kk] == 1):
                continue
            dx = self.X[x0:x1, y0:y1] - xs[kk, 0]
            dy = self.Y[x0:x1, y0:y1] - xs[kk, 1]
            d2 = dx * dx + dy * dy
            GT = G[x0:x1, y0:y1] * norm * np.exp(-d2 / (2 * w * w))
            W = pref[kk] * f[kk]
            gW[kk] = GT.sum()
            gX[kk] = W * (GT * dx).sum() / w ** 2
            gY[kk] = W * (GT * dy).sum() / w ** 2
            gw += W * ((GT * d2).sum() / w ** 3 - 2 * GT.sum() / w)
        # weights
        part = ps == 0
        dWda = np.where(part, -0.5 * f - 0.5 * (1 - alpha) * fs * 0.5 * k, f + alpha * fs * k)
        dWdk = np.where(part, 0.5 * (1 - alpha) * fs * 0.5 * (1 - alpha + 1e-8), alpha * fs * (alpha + 1e-8))
        # positions
        gxy = [np.zeros((len(self.uvw[0]), 2)), np.zeros((len(self.uvw[1]), 2))]
        kk = 0
        for nu in range(2):
            for i in self.part[nu]:
                gxy[nu][i] += (gX[kk], gY[kk])
                kk += 1
        for (i, j) in self.pairs:
            gxy[0][i] += 0.5 * np.array([gX[kk], gY[kk]])
            gxy[1][j] += 0.5 * np.array([gX[kk], gY[kk]])
            kk += 1
        gR = sum((gxy[nu] * pos[nu][1]).sum() for nu in range(2))
        ga = sum((gxy[nu] * pos[nu][2]).sum() for nu in range(2))
        return np.array([gw, gR, ga, (gW * dWda).sum(), (gW * dWdk).sum()])

    def fd_gradient(self, th, cot, h=1e-4):
        """central finite differences of cot . forward."""
        th = np.asarray(th, np.float64)
        g = np.zeros(5)
        for q in range(5):
            e = np.zeros(5)
            e[q] = h
            g[q] = (np.dot(cot, self.forward(th + e)) - np.dot(cot, self.forward(th - e))) / (2 * h)
        return g


def event_from_context(ctx, e):
    """DiffEvent for event e of a context after diff_setup (the HIP path's
    frozen topology: uniforms, b, Gamma uniforms, positions at theta0,
    participants and pairs)."""
    p, d = ctx.params, ctx.derived
    topo = ctx.diff_topology(e)
    nuc = ctx.nucleons(e)
    pa, pb, _ = ctx.pairs(e)
    g = oracle.grid(p.maxx, p.dx, p.maxy, p.dy)
    return DiffEvent(g, topo["uvw_proj"], topo["uvw_targ"], topo["b"], nuc["xy_proj"], nuc["xy_targ"],
                     nuc["wound_proj"], nuc["wound_targ"], np.stack([pa, pb], 1), topo["gamma_v"], d.width,
                     p.finalFactor, fluct=p.cc_fluctuation_model == 6)
