"""Generate the manufactured-solution fixtures of tests/golden/manufactured.npz.

    python tests/golden/make_manufactured.py

For each module, smooth analytic fields are given on the module's grid, and
sympy evaluates the module's right-hand side with EXACT derivatives:
  hydro       src/hydro/eqns.cuh:26-92        (periodic box 1 x 1)
  hydro-isot  src/hydro-isot/eqns.cuh:27-87   (periodic box 2 pi x 2 pi, with
                                               Kolmogorov forcing switched on)
  disk        src/disk/eqns.cuh:38-247        ((log r, theta) grid; R, z
                                               derivatives by sympy's chain rule
                                               from r = rin e^{log r},
                                               R = r sin(theta), z = r cos(theta))
at every cell centre, for two resolutions (n = 32 and 64).  The oracle's
discrete right-hand side must approach these values at sixth order.  The
fixture holds the field definitions' parameters, the states at both
resolutions and the exact right-hand sides (float64).
"""
import os

import numpy as np
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
X, Y = sp.symbols("x y", real=True)
TWO_PI = 2 * sp.pi


def hydro_fields():
    lnd = sp.Rational(5, 100) * sp.sin(TWO_PI * (X + 2 * Y)) + sp.Rational(3, 100) * sp.cos(TWO_PI * X)
    u1 = sp.Rational(1, 10) * sp.sin(TWO_PI * Y) + sp.Rational(5, 100) * sp.cos(TWO_PI * (X - Y))
    u2 = sp.Rational(8, 100) * sp.cos(TWO_PI * X) + sp.Rational(2, 100) * sp.sin(2 * TWO_PI * Y)
    lne = sp.log(sp.Rational(9, 10)) + sp.Rational(4, 100) * sp.sin(TWO_PI * (2 * X + Y))
    return [lnd, u1, u2, lne]


HYDRO_PAR = {"gamma": 5.0 / 3.0, "nus": 2e-3, "nub": 1e-3, "kappa": 5e-3, "dd": 1e-3}


def hydro_rhs(f, p):
    lnd, u1, u2, lne = f
    g, nu, nub, kap, dd = [sp.Float(p[k]) for k in ("gamma", "nus", "nub", "kappa", "dd")]
    dx = lambda h: sp.diff(h, X)
    dy = lambda h: sp.diff(h, Y)
    lap = lambda h: sp.diff(h, X, 2) + sp.diff(h, Y, 2)
    T = (g - 1) * sp.exp(lne)
    th = dx(u1) + dy(u2)
    s11, s12, s22 = dx(u1) - th / 3, (dx(u2) + dy(u1)) / 2, dy(u2) - th / 3
    adv = lambda h: u1 * dx(h) + u2 * dy(h)
    r_lnd = -adv(lnd) - th + dd * (lap(lnd) + dx(lnd) ** 2 + dy(lnd) ** 2)
    r_u1 = (-adv(u1) - T * (dx(lnd) + dx(lne)) + 2 * nu * (s11 * dx(lnd) + s12 * dy(lnd))
            + nu * lap(u1) + (nu / 3 + nub) * (sp.diff(u1, X, 2) + sp.diff(u2, X, Y)))
    r_u2 = (-adv(u2) - T * (dy(lnd) + dy(lne)) + 2 * nu * (s12 * dx(lnd) + s22 * dy(lnd))
            + nu * lap(u2) + (nu / 3 + nub) * (sp.diff(u1, X, Y) + sp.diff(u2, Y, 2)))
    r_lne = (-adv(lne) - (g - 1) * th
             + ((g - 1) / T) * (2 * nu * (s11 ** 2 + 2 * s12 ** 2 + s22 ** 2) + nub * th ** 2)
             + (g - 1) * kap * (lap(lne) + dx(lne) ** 2 + dy(lne) ** 2))
    return [r_lnd, r_u1, r_u2, r_lne]


def isot_fields():
    lnd = sp.Rational(6, 100) * sp.sin(X + 2 * Y) + sp.Rational(2, 100) * sp.cos(3 * X)
    u1 = sp.Rational(1, 10) * sp.sin(Y) + sp.Rational(4, 100) * sp.cos(2 * X - Y)
    u2 = sp.Rational(7, 100) * sp.cos(X) + sp.Rational(3, 100) * sp.sin(2 * Y + X)
    return [lnd, u1, u2]


ISOT_PAR = {"dd": 1e-3, "nus": 2e-3, "nub": 1e-3, "cs2": 1.3, "f": 0.05, "k": -2.0}


def isot_rhs(f, p, delta1):
    lnd, u1, u2 = f
    dd, nu, nub, cs2, ff, kk = [sp.Float(p[k]) for k in ("dd", "nus", "nub", "cs2", "f", "k")]
    dx = lambda h: sp.diff(h, X)
    dy = lambda h: sp.diff(h, Y)
    lap = lambda h: sp.diff(h, X, 2) + sp.diff(h, Y, 2)
    th = dx(u1) + dy(u2)
    s11, s12, s22 = dx(u1) - th / 3, (dx(u2) + dy(u1)) / 2, dy(u2) - th / 3
    adv = lambda h: u1 * dx(h) + u2 * dy(h)
    r_lnd = -adv(lnd) - th + dd * (lap(lnd) + dx(lnd) ** 2 + dy(lnd) ** 2)
    r_u1 = (-adv(u1) - cs2 * dx(lnd) + 2 * nu * (s11 * dx(lnd) + s12 * dy(lnd))
            + nu * lap(u1) + (nu / 3 + nub) * (sp.diff(u1, X, 2) + sp.diff(u2, X, Y)))
    # Kolmogorov forcing f cos(k i Delta1) uses the cell index i = x/Delta1 - 1/2
    r_u2 = (-adv(u2) - cs2 * dy(lnd) + 2 * nu * (s12 * dx(lnd) + s22 * dy(lnd))
            + nu * lap(u2) + (nu / 3 + nub) * (sp.diff(u1, X, Y) + sp.diff(u2, Y, 2))
            + ff * sp.cos(kk * (X - delta1 / 2)))
    return [r_lnd, r_u1, r_u2]


# disk: x is log(r / rin), y is theta
def disk_fields():
    lnd = -1 + sp.Rational(1, 10) * sp.sin(X + Y) + sp.Rational(5, 100) * sp.cos(2 * Y)
    ur = sp.Rational(5, 100) * sp.sin(Y) * sp.cos(X)
    uz = sp.Rational(4, 100) * sp.cos(Y) + sp.Rational(2, 100) * sp.sin(X - Y)
    omg = sp.Rational(2, 10) * sp.exp(-X / 2) + sp.Rational(1, 100) * sp.cos(Y)
    lne = sp.log(sp.Rational(3, 10)) + sp.Rational(1, 10) * sp.cos(X + 2 * Y)
    lner = sp.log(sp.Rational(1, 10)) + sp.Rational(1, 10) * sp.sin(2 * X - Y)
    return [lnd, ur, uz, omg, lne, lner]


DISK_PAR = {"M": 1.0, "rS": 0.5, "rin": 3.0, "gamma": 5.0 / 3.0, "nus": 1e-3, "nub": 5e-4,
            "kappa": 2e-3, "alpha": 0.01, "ar": 2.0, "es": 1.0, "ff": 0.05, "dd": 1e-4,
            "nu": 2e-4, "ed": 5e-4, "rd": 5e-4}


def disk_rhs(f, p):
    lnd, ur, uz, Omg, lne, lner = f
    P = {k: sp.Float(v) for k, v in p.items()}
    sphr = P["rin"] * sp.exp(X)
    sint, cost = sp.sin(Y), sp.cos(Y)
    r = sphr * sint
    # physical derivatives by the chain rule of (R, z) = r (sin, cos)(theta)
    dR = lambda h: (sint * sp.diff(h, X) + cost * sp.diff(h, Y)) / sphr
    dZ = lambda h: (cost * sp.diff(h, X) - sint * sp.diff(h, Y)) / sphr
    lap = lambda h: (sp.diff(h, X, 2) + sp.diff(h, Y, 2)) / sphr ** 2   # module's lap
    den, eng, rad = sp.exp(lnd), sp.exp(lne), sp.exp(lner)
    g = P["gamma"]
    g1 = g - 1
    temp = g1 * eng
    nuSS = P["alpha"] * g * temp * r * sp.sqrt(r / P["M"])
    nut = P["nus"] + nuSS
    out = {}
    out["lnd"] = -(ur * dR(lnd) + uz * dZ(lnd))
    out["ur"] = -(ur * dR(ur) + uz * dZ(ur) - Omg ** 2 * r)
    out["uz"] = -(ur * dR(uz) + uz * dZ(uz))
    out["Omg"] = -(ur * dR(Omg) + uz * dZ(Omg) + 2 * ur * Omg / r)
    out["lne"] = -(ur * dR(lne) + uz * dZ(lne))
    out["lner"] = -(ur * dR(lner) + uz * dZ(lner))
    ur_r = ur / r
    div_u = dR(ur) + dZ(uz) + ur_r
    out["lnd"] -= div_u
    out["ur"] -= temp * (dR(lnd) + dR(lne))
    out["uz"] -= temp * (dZ(lnd) + dZ(lne))
    out["lne"] -= div_u * g1
    Srr, Szz, Spp = dR(ur) - div_u / 3, dZ(uz) - div_u / 3, ur_r - div_u / 3
    Srz = (dZ(ur) + dR(uz)) / 2
    Szp, Spr = r * dZ(Omg) / 2, r * dR(Omg) / 2
    two = 2 * nut
    out["ur"] += two * (Srr * dR(lnd) + Srz * dZ(lnd))
    out["uz"] += two * (Srz * dR(lnd) + Szz * dZ(lnd))
    out["Omg"] += two * (Spr * dR(lnd) + Szp * dZ(lnd)) / r
    out["lne"] += (two * (Srr ** 2 + Szz ** 2 + Spp ** 2 + 2 * (Srz ** 2 + Szp ** 2 + Spr ** 2))
                   + P["nub"] * div_u ** 2) / eng
    tmp1 = nut + P["nu"] * sphr ** 2
    tmp2 = nut / r
    out["ur"] += tmp1 * lap(ur) + tmp2 * (dR(ur) - ur / r)
    out["uz"] += tmp1 * lap(uz) + tmp2 * dR(uz)
    out["Omg"] += tmp1 * lap(Omg) + tmp2 * (3 * dR(Omg))
    mixed = nut / 3 + P["nub"]
    out["ur"] += mixed * (dR(dR(ur)) + dR(dZ(uz)) + (dR(ur) - ur / r) / r)
    out["uz"] += mixed * (dR(dZ(ur)) + dZ(dZ(uz)) + dZ(ur) / r)
    lap_d_d = lap(lnd) + dR(lnd) ** 2 + dZ(lnd) ** 2
    lap_e_e = lap(lne) + dR(lne) ** 2 + dZ(lne) ** 2
    out["lnd"] += lap_d_d * P["dd"] * sphr ** 2
    out["lne"] += lap_e_e * (P["ed"] * sphr ** 2 + P["kappa"] * g1 + g * nuSS)
    kappa = P["ff"] * den ** 2 * temp ** sp.Float(-3.5)
    chi = P["es"] * den + kappa
    d_lner_2 = dR(lner) ** 2 + dZ(lner) ** 2
    bigR = sp.sqrt(d_lner_2) / chi
    denom = 6 + (3 + bigR) * bigR
    Lam = (2 + bigR) / denom
    out["ur"] -= Lam * rad * (dR(lnd) + dR(lne))
    out["uz"] -= Lam * rad * (dZ(lnd) + dZ(lne))
    cooling = kappa * (P["ar"] * temp ** 4 / den - rad)
    out["lne"] -= cooling / eng
    out["lner"] += cooling / rad
    lnEr = lnd + lner
    d_lnd_d_lner = dR(lnd) * dR(lner) + dZ(lnd) * dZ(lner)
    lap_er_er = lap(lner) + d_lner_2
    nu_rad = Lam / chi
    Lamp = -bigR * (4 + bigR) / denom ** 2
    dr_chi = chi * dR(lnd) - sp.Float(3.5) * kappa * dR(lne)
    dz_chi = chi * dZ(lnd) - sp.Float(3.5) * kappa * dZ(lne)
    eps = sp.Float(1e-16)
    dr_bigR = ((dR(lnEr) * dR(dR(lnEr)) + dZ(lnEr) * dR(dZ(lnEr))) / (eps + bigR) / chi
               - bigR * dr_chi) / chi
    dz_bigR = ((dR(lnEr) * dR(dZ(lnEr)) + dZ(lnEr) * dZ(dZ(lnEr))) / (eps + bigR) / chi
               - bigR * dz_chi) / chi
    out["lner"] += (lap_er_er * P["rd"] * sphr ** 2
                    + nu_rad * (lap_d_d + 2 * d_lnd_d_lner + lap_er_er)
                    + dR(lnEr) * (Lamp * dr_bigR - nu_rad * dr_chi) / chi
                    + dZ(lnEr) * (Lamp * dz_bigR - nu_rad * dz_chi) / chi)
    gr = P["M"] / (sphr - P["rS"]) ** 2
    out["ur"] -= sint * gr
    out["uz"] -= cost * gr
    return [out[k] for k in ("lnd", "ur", "uz", "Omg", "lne", "lner")]


def grid(n1, n2, l1, l2):
    x = l1 * (np.arange(n1) + 0.5) / n1
    y = l2 * (np.arange(n2) + 0.5) / n2
    return np.meshgrid(x, y, indexing="ij")


def evaluate(exprs, Xg, Yg):
    out = []
    for e in exprs:
        f = sp.lambdify((X, Y), e, "numpy")
        v = np.asarray(f(Xg, Yg), dtype=np.float64)
        out.append(np.broadcast_to(v, Xg.shape).copy())
    return np.stack(out, axis=-1)


def main():
    data = {}
    for n in (32, 64):
        # hydro: the default box l1 = l2 = 1 (src/global.cpp:28)
        Xg, Yg = grid(n, n, 1.0, 1.0)
        f = hydro_fields()
        data[f"hydro_{n}_state"] = evaluate(f, Xg, Yg)
        data[f"hydro_{n}_rhs"] = evaluate(hydro_rhs(f, HYDRO_PAR), Xg, Yg)
        # hydro-isot: box 2 pi (src/hydro-isot/config.h:25), in FP64 as the
        # FP64 oracle holds it
        L = 2 * np.pi
        Xg, Yg = grid(n, n, L, L)
        f = isot_fields()
        d1 = L / n
        data[f"isot_{n}_state"] = evaluate(f, Xg, Yg)
        data[f"isot_{n}_rhs"] = evaluate(isot_rhs(f, ISOT_PAR, d1), Xg, Yg)
        # disk: l2 = pi, l1 = 2 asinh(pi / 2n) n (src/disk/config.h:21-34), FP64
        l2 = np.pi
        hd2 = 0.5 * l2 / n
        hd1 = np.log(hd2 + np.sqrt(hd2 * hd2 + 1.0))
        l1 = 2.0 * hd1 * n
        Xg, Yg = grid(n, n, l1, l2)
        f = disk_fields()
        data[f"disk_{n}_state"] = evaluate(f, Xg, Yg)
        data[f"disk_{n}_rhs"] = evaluate(disk_rhs(f, DISK_PAR), Xg, Yg)
        print("done", n, flush=True)
    np.savez_compressed(os.path.join(HERE, "manufactured.npz"), **data)


if __name__ == "__main__":
    main()
