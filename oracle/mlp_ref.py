"""CPU oracle of the learned interface solver (numpy).

Test infrastructure only: imported by tests/ and by the cpu_baseline leg of bench.py as the
checker; the product package never imports it.

* ``forward`` / ``loss_grad``: the 10-256-256-256-256-3 SiLU network and its back-propagated
  gradient, either in exact float64 (``emulate=False``, pinned against central finite differences
  in the tests) or with the GPU kernels' rounding points (``emulate=True``): bf16 inputs, weights,
  post-activations, SiLU' and back-propagated dZ, FP32 sums and biases.
* ``features`` / ``targets``: the input and output maps of the network (stiffened-gas pairs).
* ``sg_star`` / ``sg_sample0``: the closed-form stiffened-gas Riemann solution (Toro, "Riemann
  Solvers and Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4, with p -> p + p_inf), pinned in
  the tests against Toro's Table 4.2 and against the oracle's restatement of
  ExactRiemannSolverBase::ComputeRiemannSolution (ExactRiemannSolverBase.cpp:59-342).
"""
import numpy as np

NIN, NH, NOUT = 10, 256, 3
DIMS = [(NH, NIN), (NH, NH), (NH, NH), (NH, NH), (NOUT, NH)]


def offsets():
    """(weight offset, bias offset) of each layer in the flat parameter vector."""
    out, o = [], 0
    for (no, ni) in DIMS:
        out.append((o, o + no * ni))
        o += no * ni + no
    return out, o


OFFS, NPARAM = offsets()


def unpack(p):
    W, b = [], []
    for (wo, bo), (no, ni) in zip(OFFS, DIMS):
        W.append(np.asarray(p[wo:wo + no * ni], dtype=np.float64).reshape(no, ni))
        b.append(np.asarray(p[bo:bo + no], dtype=np.float64))
    return W, b


def pack(W, b):
    p = np.empty(NPARAM, dtype=np.float64)
    for (wo, bo), w, bb in zip(OFFS, W, b):
        p[wo:wo + w.size] = w.ravel()
        p[bo:bo + bb.size] = bb
    return p


def bf16(x):
    """Round to bfloat16 (round to nearest, ties to even) through float32; returns float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).astype(np.float64)
    return np.where(np.isnan(f), np.nan, out)


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s * (1.0 + z * (1.0 - s))


def forward(p, X, emulate=False):
    """Network outputs Y [n][3] and the activations the backward pass needs."""
    W, b = unpack(p)
    rd = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    fs = f32 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    H = [rd(X)]
    D = [None]
    for l in range(4):
        Z = fs(H[l] @ rd(W[l]).T + fs(b[l]))
        h, d = silu(Z)
        H.append(rd(h))
        D.append(rd(d))
    Y = fs(H[4] @ rd(W[4]).T + fs(b[4]))
    return Y, (H, D)


def loss_grad(p, X, T, emulate=False):
    """loss = sum over rows/outputs of (y - t)^2 / n (rows with NaN targets excluded); gradient (flat)."""
    n = X.shape[0]
    W, b = unpack(p)
    rd = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    fs = f32 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    Y, (H, D) = forward(p, X, emulate)
    mask = ~np.isnan(T)
    d = np.where(mask, Y - np.nan_to_num(T), 0.0)
    loss = float(np.sum(d * d) / n)
    dZ = rd(fs(2.0 / n * d))                       # gradient at the output layer
    gW, gb = [None] * 5, [None] * 5
    for l in range(4, -1, -1):
        gW[l] = dZ.T @ H[l]
        gb[l] = dZ.sum(axis=0)
        if l > 0:
            dZ = rd(fs(dZ @ rd(W[l])) * D[l])
    return loss, pack(gW, gb)


# ---------------------------------------------------------------- stiffened-gas Riemann problem

class Side:
    def __init__(self, rho, u, p, gam, pinf):
        self.rho, self.u, self.p, self.gam, self.pinf = float(rho), float(u), float(p), float(gam), float(pinf)
        self.P = self.p + self.pinf
        self.c = np.sqrt(self.gam * self.P / self.rho)


def wave(s, p):
    """f_K(p), df_K/dp (Toro 4.6-4.7 in effective pressure P = p + p_inf)."""
    Pp = p + s.pinf
    if p > s.p:
        A = 2.0 / ((s.gam + 1.0) * s.rho)
        B = (s.gam - 1.0) / (s.gam + 1.0) * s.P
        q = np.sqrt(A / (Pp + B))
        return (p - s.p) * q, q * (1.0 - 0.5 * (p - s.p) / (Pp + B))
    r = Pp / s.P
    rw = r ** ((s.gam - 1.0) / (2.0 * s.gam))
    return 2.0 * s.c / (s.gam - 1.0) * (rw - 1.0), rw / (r * s.rho * s.c)


def star_rho(s, p):
    r = (p + s.pinf) / s.P
    if p > s.p:
        mu = (s.gam - 1.0) / (s.gam + 1.0)
        return s.rho * (r + mu) / (mu * r + 1.0)
    return s.rho * r ** (1.0 / s.gam)


def residual(L, R, p):
    """Velocity continuity residual |u*_L(p) - u*_R(p)| / max(c_L, c_R)."""
    fl, _ = wave(L, p)
    fr, _ = wave(R, p)
    return abs(fl + fr + R.u - L.u) / max(L.c, R.c)


def newton(L, R, p0, nmax, tol):
    """(accepted, p, updates): at most nmax Newton updates from p0 until residual <= tol, then one
    polishing update (kept when it does not raise the residual)."""
    pfloor = -min(L.pinf, R.pinf)
    p = p0 if p0 > pfloor else pfloor + 1e-6 * min(L.P, R.P)
    it = 0
    met, pok, rok = False, 0.0, 0.0
    while True:
        fl, dl = wave(L, p)
        fr, dr = wave(R, p)
        F = fl + fr + R.u - L.u
        res = abs(F) / max(L.c, R.c)
        if met:
            return True, (p if res <= rok else pok), it
        if res <= tol:
            met, pok, rok = True, p, res
        elif it == nmax or not np.isfinite(res):
            return False, p, it
        pn = p - F / (dl + dr)
        if not pn > pfloor:
            pn = 0.5 * (p + pfloor)
        if met and not np.isfinite(pn):
            return True, pok, it
        p = pn
        it += 1


def sg_star(L, R):
    """Closed-form star state (p*, u*, rho*_L, rho*_R) to machine precision (bisection-safeguarded
    Newton on F)."""
    pfloor = -min(L.pinf, R.pinf)
    lo, hi = pfloor, max(L.p, R.p)
    F = lambda p: wave(L, p)[0] + wave(R, p)[0] + R.u - L.u
    while F(hi) < 0.0:
        hi = pfloor + 2.0 * (hi - pfloor)
    p = 0.5 * (lo + hi)
    for _ in range(400):
        fl, dl = wave(L, p)
        fr, dr = wave(R, p)
        f = fl + fr + R.u - L.u
        if f > 0:
            hi = p
        else:
            lo = p
        pn = p - f / (dl + dr)
        if not (lo < pn < hi):
            pn = 0.5 * (lo + hi)
        if pn == p:
            break
        p = pn
    fl, _ = wave(L, p)
    fr, _ = wave(R, p)
    us = 0.5 * (L.u + R.u) + 0.5 * (fr - fl)
    return p, us, star_rho(L, p), star_rho(R, p)


def sg_sample0(L, R, ps, us, rl, rr):
    """(rho, u, p) at x/t = 0 (Toro sec. 4.5; fans 4.56, 4.63)."""
    if us >= 0.0:
        Ps = ps + L.pinf
        if ps > L.p:
            SL = L.u - L.c * np.sqrt((L.gam + 1) / (2 * L.gam) * Ps / L.P + (L.gam - 1) / (2 * L.gam))
            return (L.rho, L.u, L.p) if SL >= 0 else (rl, us, ps)
        if L.u - L.c >= 0:
            return (L.rho, L.u, L.p)
        if us - L.c * (Ps / L.P) ** ((L.gam - 1) / (2 * L.gam)) <= 0:
            return (rl, us, ps)
        u = 2.0 / (L.gam + 1) * (L.c + 0.5 * (L.gam - 1) * L.u)
        q = u / L.c
        return (L.rho * q ** (2 / (L.gam - 1)), u, L.P * q ** (2 * L.gam / (L.gam - 1)) - L.pinf)
    Ps = ps + R.pinf
    if ps > R.p:
        SR = R.u + R.c * np.sqrt((R.gam + 1) / (2 * R.gam) * Ps / R.P + (R.gam - 1) / (2 * R.gam))
        return (R.rho, R.u, R.p) if SR <= 0 else (rr, us, ps)
    if R.u + R.c <= 0:
        return (R.rho, R.u, R.p)
    if us + R.c * (Ps / R.P) ** ((R.gam - 1) / (2 * R.gam)) >= 0:
        return (rr, us, ps)
    u = 2.0 / (R.gam + 1) * (-R.c + 0.5 * (R.gam - 1) * R.u)
    q = -u / R.c
    return (R.rho * q ** (2 / (R.gam - 1)), u, R.P * q ** (2 * R.gam / (R.gam - 1)) - R.pinf)


def features(L, R):
    """The 10 network inputs (float32) and the scales (Ps, pimin, cs, ubar)."""
    cs = L.c + R.c
    ZL, ZR = L.rho * L.c, R.rho * R.c
    pimin = min(L.pinf, R.pinf)
    Ps = np.sqrt(L.P * R.P)
    du = R.u - L.u
    pac = (ZR * L.p + ZL * R.p - ZL * ZR * du) / (ZL + ZR)
    qac = max((pac + pimin) / Ps, 1e-8)
    x = np.array([np.log(R.rho / L.rho) / 7.0, np.log(R.P / L.P) / 9.0, du / cs, np.log(R.c / L.c) / 5.0,
                  np.log(L.gam), np.log(R.gam), L.pinf / L.P, R.pinf / R.P, np.log(ZR / ZL) / 7.0,
                  np.log(qac) * 0.25], dtype=np.float32)
    return x, (Ps, pimin, cs, 0.5 * (L.u + R.u))


def targets(L, R, ps, us, rl):
    """The 3 network outputs of a star state."""
    _, (Ps, pimin, cs, ubar) = features(L, R)
    return np.array([0.25 * np.log((ps + pimin) / Ps), (us - ubar) / cs, 0.25 * np.log(rl / L.rho)], dtype=np.float32)


def decode_p(x_scales, y0):
    Ps, pimin, _, _ = x_scales
    return np.exp(4.0 * float(y0)) * Ps - pimin
