"""Generate the piecewise-polynomial Rys root/weight tables (rys_tables.inc).

Mathematics (not taken from any code): for the Rys weight exp(-x t^2) on
t in [0, 1], substitute s = t^2.  The n-point Gauss rule of the measure
    d mu_x(s) = s^(-1/2) exp(-x s) ds / 2   on [0, 1]
has ordinary moments mu_k = F_k(x) (the Boys function), so its nodes s_i and
weights w_i satisfy  sum_i w_i s_i^k = F_k(x)  for k < 2n.  libcint's ERI
recurrences consume u_i = s_i / (1 - s_i) and w_i (src/g2e.c:4518-4540).

Procedure, at 40 significant digits with mpmath:
  1. F_k(x) = gamma_lower(k + 1/2, x) / (2 x^(k + 1/2)),  F_k(0) = 1/(2k+1).
  2. Gautschi's Chebyshev algorithm turns the 2n moments into the three-term
     recurrence coefficients (alpha_k, beta_k).
  3. Golub-Welsch: eigenvalues of the symmetric Jacobi matrix are s_i, the
     squared first eigenvector components times mu_0 are w_i.
For x >= X_n the Rys rule equals, to below 1e-16 relative, the rule of the
measure s^(-1/2) e^(-s) ds / 2 on [0, inf) rescaled by x:  u_i = r_i/(x-r_i),
w_i = W_i / (2 sqrt(x)), where (r_i, W_i) is the n-point generalized
Gauss-Laguerre rule with alpha = -1/2.  X_n = 36 + 6 n was found by scanning
the difference between the two rules (it drops below 1e-16 near 34 + 6 n).

Below X_n, each unit interval [a, a+1) carries, for every root and weight,
the degree-13 polynomial in z = 2 (x - a) - 1 that interpolates the function
at the 14 Chebyshev points of the interval.  Coefficients are written in the
monomial basis of z, so a lane evaluates each function with 13 FMAs (Horner).

Output: ../csrc/rys_tables.inc (C source, included by the HIP kernels and by
the CPU restatement's fast path).  Run:  python gen_rys_tables.py [--jobs 8]
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from multiprocessing import Pool

import mpmath as mp

NMAX = 9
DEG = 13
DPS = 40


def xmax(n: int) -> int:
    return 36 + 6 * n


def moments(n, x):
    x = mp.mpf(x)
    half = mp.mpf(1) / 2
    if x == 0:
        return [mp.mpf(1) / (2 * k + 1) for k in range(2 * n)]
    return [mp.gammainc(k + half, 0, x) / (2 * x ** (k + half)) for k in range(2 * n)]


def gauss_from_moments(mu, n):
    """Chebyshev algorithm + Golub-Welsch; returns ascending nodes and weights."""
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mu)
    alpha = [mu[1] / mu[0]]
    beta = [mu[0]]
    for k in range(1, n):
        sig_new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            sig_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha.append(sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1])
        beta.append(sig_new[k] / sig[k - 1])
        sig_prev, sig = sig, sig_new
    jac = mp.matrix(n, n)
    for i in range(n):
        jac[i, i] = alpha[i]
        if i + 1 < n:
            jac[i, i + 1] = jac[i + 1, i] = mp.sqrt(beta[i + 1])
    ev, vec = mp.eigsy(jac)
    order = sorted(range(n), key=lambda i: ev[i])
    return [ev[i] for i in order], [mu[0] * vec[0, i] ** 2 for i in order]


def rys_u_w(n, x):
    s, w = gauss_from_moments(moments(n, x), n)
    return [si / (1 - si) for si in s], w


def laguerre_half(n):
    """n-point Gauss rule for s^(-1/2) e^(-s) on [0, inf): alpha_k=2k+1/2, beta_k=k(k-1/2)."""
    jac = mp.matrix(n, n)
    half = mp.mpf(1) / 2
    for k in range(n):
        jac[k, k] = 2 * k + half
        if k + 1 < n:
            jac[k, k + 1] = jac[k + 1, k] = mp.sqrt((k + 1) * (k + 1 - half))
    ev, vec = mp.eigsy(jac)
    order = sorted(range(n), key=lambda i: ev[i])
    return [ev[i] for i in order], [mp.sqrt(mp.pi) * vec[0, i] ** 2 for i in order]


def _cheb_to_monomial(c):
    """Chebyshev-series coefficients -> monomial coefficients (exact in mp)."""
    deg = len(c) - 1
    # T_j as monomial coefficient lists
    t_prev = [mp.mpf(1)] + [mp.mpf(0)] * deg
    t_cur = [mp.mpf(0), mp.mpf(1)] + [mp.mpf(0)] * (deg - 1)
    out = [mp.mpf(0)] * (deg + 1)
    for k in range(deg + 1):
        out[k] += c[0] * t_prev[k]
    if deg >= 1:
        for k in range(deg + 1):
            out[k] += c[1] * t_cur[k]
    for j in range(2, deg + 1):
        t_next = [mp.mpf(0)] * (deg + 1)
        for k in range(deg + 1):
            if k >= 1:
                t_next[k] += 2 * t_cur[k - 1]
            t_next[k] -= t_prev[k]
        for k in range(deg + 1):
            out[k] += c[j] * t_next[k]
        t_prev, t_cur = t_cur, t_next
    return out


def fit_interval(args):
    """Monomial-in-z coefficients [DEG+1][2n] for x in [a, a+1)."""
    n, a = args
    mp.mp.dps = DPS
    npts = DEG + 1
    zs = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / npts) for k in range(npts)]
    vals = []
    for z in zs:
        u, w = rys_u_w(n, mp.mpf(a) + (z + 1) / 2)
        vals.append(u + w)
    coefs = []
    for f in range(2 * n):
        c = []
        for j in range(npts):
            acc = mp.mpf(0)
            for k in range(npts):
                acc += vals[k][f] * mp.cos(mp.pi * j * (k + mp.mpf(1) / 2) / npts)
            c.append(acc * 2 / npts)
        c[0] /= 2
        coefs.append(_cheb_to_monomial(c))
    # [j][f]
    return [[float(coefs[f][j]) for f in range(2 * n)] for j in range(npts)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "..", "csrc", "rys_tables.inc"))
    args = ap.parse_args()
    mp.mp.dps = DPS
    t0 = time.time()
    jobs = [(n, a) for n in range(1, NMAX + 1) for a in range(xmax(n))]
    with Pool(args.jobs) as pool:
        res = pool.map(fit_interval, jobs, chunksize=4)
    table = {}
    for (n, a), r in zip(jobs, res):
        table.setdefault(n, []).append(r)

    lines = []
    lines.append("/* Generated by libcint_amd/tables/gen_rys_tables.py -- do not edit.")
    lines.append(" * Rys roots u_i = t_i^2/(1-t_i^2) and weights w_i for the weight exp(-x t^2)")
    lines.append(" * on [0,1], n = 1..%d.  For x < RYS_XMAX(n): unit intervals, degree-%d" % (NMAX, DEG))
    lines.append(" * polynomials in z = 2(x - floor(x)) - 1, layout [interval][coef][root..., weight...].")
    lines.append(" * For x >= RYS_XMAX(n): u_i = r_i/(x - r_i), w_i = rys_lag_w[i]/sqrt(x). */")
    lines.append("#define RYS_NMAX %d" % NMAX)
    lines.append("#define RYS_NCOEF %d" % (DEG + 1))
    lines.append("#define RYS_XMAX(n) (36 + 6 * (n))")
    offs = []
    off = 0
    for n in range(1, NMAX + 1):
        offs.append(off)
        off += xmax(n) * (DEG + 1) * 2 * n
    lines.append("#define RYS_TABLE_SIZE %d" % off)
    lines.append("RYS_TABLE_QUAL int rys_table_offset[%d] = {0, %s};" % (NMAX + 1, ", ".join(str(o) for o in offs)))
    lag_off = []
    lo = 0
    lag_r, lag_w = [], []
    for n in range(1, NMAX + 1):
        r, w = laguerre_half(n)
        lag_off.append(lo)
        lo += n
        lag_r.extend(r)
        lag_w.extend([wi / 2 for wi in w])
    lines.append("RYS_TABLE_QUAL int rys_lag_offset[%d] = {0, %s};" % (NMAX + 1, ", ".join(str(o) for o in lag_off)))
    lines.append("RYS_TABLE_QUAL double rys_lag_r[%d] = {" % lo)
    lines.extend("  %s," % mp.nstr(v, 20, min_fixed=-100, max_fixed=100) for v in lag_r)
    lines.append("};")
    lines.append("RYS_TABLE_QUAL double rys_lag_w[%d] = {" % lo)
    lines.extend("  %s," % mp.nstr(v, 20, min_fixed=-100, max_fixed=100) for v in lag_w)
    lines.append("};")
    lines.append("RYS_TABLE_QUAL double rys_coef[RYS_TABLE_SIZE] = {")
    for n in range(1, NMAX + 1):
        lines.append("/* n = %d */" % n)
        for a, blk in enumerate(table[n]):
            for j, row in enumerate(blk):
                lines.append("  " + ", ".join(repr(v) for v in row) + ",")
    lines.append("};")
    with open(args.out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("wrote %s (%d doubles) in %.1fs" % (args.out, off, time.time() - t0), file=sys.stderr)


if __name__ == "__main__":
    main()
