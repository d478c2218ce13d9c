/*
 * CPU restatement of libcint's two-electron Rys-quadrature path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the GPU
 * library: only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark load it, always as the checker and never as the product.
 *
 * It restates, in plain C, the algorithm of the reference library
 * (libcint 6.1.1), function by function:
 *   env set-up             src/g2e.c:21-172          (oc_init_2e)
 *   3c2e env set-up        src/g3c2e.c:21-132        (oc_init_3c2e)
 *   pair data + screening  src/optimizer.c:288-342   (oc_pairdata)
 *   primitive loop         src/cint2e.c:57-269       (oc_loop_2e)
 *   3c2e primitive loop    src/cint3c2e.c:56-211     (oc_loop_3c2e)
 *   Rys coefficients       src/g2e.c:4425-4545       (oc_g0_2e)
 *   2D VRR                 src/g2e.c:272-421         (oc_vrr)
 *   HRR (2d -> 4d)         src/g2e.c:428-693         (oc_hrr)
 *   gout (ERI)             src/cint2e.c:961-1183     (oc_gout_plain)
 *   gout (ip1) + nabla_i   src/autocode/grad2.c:19-68, src/g2e.c:4550-4591
 *   cart -> sph            src/cart2sph.c:5324-5376, 5884-5924 (oc_c2s_*)
 *   common factor          src/g1e.c:565-572
 *   radial normalisation   src/misc.c:72-93
 *
 * Rys roots and weights come from one of two modes:
 *   ORC_ROOTS_EXACT  the Gauss rule of exp(-x t^2) on [0,1] built from the
 *                    Boys-function moments in __float128 (Chebyshev
 *                    algorithm + implicit-shift QL, i.e. Golub-Welsch), the
 *                    same mathematics the reference reaches with its
 *                    polynomial fits (src/rys_roots.c:57-122) and its
 *                    Wheeler/Schmidt solvers (src/rys_wheeler.c:3667-3715,
 *                    src/rys_roots.c:1645-1769);
 *   ORC_ROOTS_TABLE  the generated piecewise polynomials of
 *                    libcint_amd/csrc/rys_tables.inc, evaluated here by
 *                    this file's own code (fast; checked against EXACT).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <quadmath.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RYS_TABLE_QUAL static const
#include "../libcint_amd/csrc/rys_tables.inc"

/* env / atm / bas slots, include/cint.h.in:26-67 */
#define PTR_EXPCUTOFF 0
#define PTR_RANGE_OMEGA 8
#define ATOM_OF 0
#define ANG_OF 1
#define NPRIM_OF 2
#define NCTR_OF 3
#define PTR_EXP 5
#define PTR_COEFF 6
#define BAS_SLOTS 8
#define PTR_COORD 1
#define ATM_SLOTS 6
#define BAS(s, i) bas[BAS_SLOTS * (i) + (s)]
#define ATM(s, i) atm[ATM_SLOTS * (i) + (s)]

#define OC_EXPCUTOFF 60.0     /* src/cint_config.h.in:27 */
#define OC_MIN_EXPCUTOFF 40.0 /* src/cint_config.h.in:30 */
#define OC_MAXROOTS 16
#define OC_EXPCUTOFF_SR 40.0  /* src/rys_roots.h:46 */
#define OC_PI 3.1415926535897932384626433832795028
#define OC_SQRTPI 1.7724538509055160272981674833411451

enum { ORC_ROOTS_EXACT = 0, ORC_ROOTS_TABLE = 1 };
static int g_root_mode = ORC_ROOTS_EXACT;

void orc_set_root_mode(int mode) { g_root_mode = mode; }
int orc_get_root_mode(void) { return g_root_mode; }

/* ------------------------------------------------------------------ */
/* Boys function F_m(x), m = 0..mmax, in quad precision.               */
/* Series for the top order, then the downward recurrence              */
/*   F_{m-1} = (2 x F_m + e^-x) / (2m - 1).                            */
/* ------------------------------------------------------------------ */
static void boys_q(int mmax, __float128 x, __float128 *f)
{
        __float128 ex = expq(-x);
        if (x > 200) {
                /* upward recurrence from F_0 is stable for x >> m */
                __float128 sx = sqrtq(x);
                f[0] = sqrtq(M_PIq) * 0.5Q * erfq(sx) / sx;
                for (int m = 0; m < mmax; m++)
                        f[m + 1] = ((2 * m + 1) * f[m] - ex) / (2 * x);
                return;
        }
        __float128 term = 1.0Q / (2 * mmax + 1);
        __float128 sum = term;
        for (int k = 1; k < 2000; k++) {
                term *= 2 * x / (2 * mmax + 2 * k + 1);
                sum += term;
                if (term < sum * 1e-36Q)
                        break;
        }
        f[mmax] = ex * sum;
        for (int m = mmax; m > 0; m--)
                f[m - 1] = (2 * x * f[m] + ex) / (2 * m - 1);
}

void orc_boys(int mmax, double x, double *out)
{
        __float128 f[64];
        if (mmax > 63) mmax = 63;
        boys_q(mmax, (__float128)x, f);
        for (int m = 0; m <= mmax; m++)
                out[m] = (double)f[m];
}

/* ------------------------------------------------------------------ */
/* Exact Rys rule: moments -> recurrence (Chebyshev algorithm) ->       */
/* Jacobi matrix eigen-decomposition (implicit QL, first eigenvector    */
/* row only).                                                           */
/* ------------------------------------------------------------------ */
static void ql_first_row(int n, __float128 *d, __float128 *e, __float128 *z)
{
        for (int i = 0; i < n; i++) z[i] = (i == 0);
        e[n - 1] = 0;
        for (int l = 0; l < n; l++) {
                int iter = 0;
                int m;
                for (;;) {
                        for (m = l; m < n - 1; m++) {
                                __float128 dd = fabsq(d[m]) + fabsq(d[m + 1]);
                                if (fabsq(e[m]) <= 1e-33Q * dd) break;
                        }
                        if (m == l) break;
                        if (++iter > 200) break;
                        __float128 g = (d[l + 1] - d[l]) / (2 * e[l]);
                        __float128 r = hypotq(g, 1.0Q);
                        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
                        __float128 s = 1, c = 1, p = 0;
                        int i;
                        int deflated = 0;
                        for (i = m - 1; i >= l; i--) {
                                __float128 f = s * e[i];
                                __float128 b = c * e[i];
                                r = hypotq(f, g);
                                e[i + 1] = r;
                                if (r == 0) {
                                        d[i + 1] -= p;
                                        e[m] = 0;
                                        deflated = 1;
                                        break;
                                }
                                s = f / r;
                                c = g / r;
                                g = d[i + 1] - p;
                                r = (d[i] - g) * s + 2 * c * b;
                                p = s * r;
                                d[i + 1] = g + p;
                                g = c * r - b;
                                __float128 zt = z[i + 1];
                                z[i + 1] = s * z[i] + c * zt;
                                z[i] = c * z[i] - s * zt;
                        }
                        if (deflated) continue;
                        d[l] -= p;
                        e[l] = g;
                        e[m] = 0;
                }
        }
}

/* Gauss rule from the moments mu[k] = int t^2k W(t) dt, k < 2n:
 * Chebyshev algorithm -> Jacobi matrix -> QL; nodes s = t^2 ascending. */
static void gauss_from_moments_q(int n, const __float128 *mu, __float128 *s_out, __float128 *w_out);

static void rys_exact_q(int n, double x, __float128 *s_out, __float128 *w_out)
{
        __float128 mu[2 * OC_MAXROOTS];
        boys_q(2 * n - 1, (__float128)x, mu);
        gauss_from_moments_q(n, mu, s_out, w_out);
}

/* Short-range rule: the n-point Gauss rule of exp(-x t^2) on [lower, 1]
 * in the variable s = t^2 (the rule CINTsr_rys_roots returns,
 * src/rys_roots.c:149-267; its moments are fmt_erfc_like's,
 * src/fmt.c:340-379).  Ordinary moments are hopelessly ill-conditioned for
 * narrow [lower, 1], so the rule is built by the discretised Stieltjes
 * procedure: the weight is sampled by 16-point Gauss-Legendre on
 * subintervals across which exp(-x t^2) changes by at most e^2 (up to where
 * it has fallen by e^-100), the three-term recurrence of that discrete
 * measure is generated in long double, and its Jacobi matrix is
 * diagonalised like the full-range rule's. */
static void gauss_legendre16(long double *xg, long double *wg)
{
        const int m = 16;
        for (int i = 0; i < m; i++) {
                long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (m + 0.5L));
                for (int it = 0; it < 100; it++) {
                        long double p0 = 1, p1 = z;
                        for (int k = 2; k <= m; k++) {
                                long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                                p0 = p1;
                                p1 = p2;
                        }
                        long double dp = m * (z * p1 - p0) / (z * z - 1);
                        long double dz = p1 / dp;
                        z -= dz;
                        if (fabsl(dz) < 1e-19L) {
                                xg[i] = z;
                                wg[i] = 2 / ((1 - z * z) * dp * dp);
                                break;
                        }
                }
        }
}

#define SR_MAX_POINTS (16 * 64)

static void rys_sr_rule(int n, double x, double lower, double *u, double *w)
{
        long double xg[16], wg[16];
        gauss_legendre16(xg, wg);
        long double lo2 = (long double)lower * lower;
        long double s_end = 1;
        if (x > 0 && lo2 + 100.0L / x < 1) s_end = lo2 + 100.0L / x;
        int K = (int)ceill((s_end - lo2) * x / 2);
        if (K < 4) K = 4;
        if (K > SR_MAX_POINTS / 16) K = SR_MAX_POINTS / 16;
        static __thread long double ps[SR_MAX_POINTS], pw[SR_MAX_POINTS];
        int M = 0;
        for (int k = 0; k < K; k++) {
                long double t0 = sqrtl(lo2 + (s_end - lo2) * k / K);
                long double t1 = sqrtl(lo2 + (s_end - lo2) * (k + 1) / K);
                long double mid = (t0 + t1) / 2, half = (t1 - t0) / 2;
                for (int g = 0; g < 16; g++) {
                        long double t = mid + half * xg[g];
                        ps[M] = t * t;
                        pw[M] = half * wg[g] * expl(-(long double)x * t * t);
                        M++;
                }
        }
        /* Stieltjes: p_{k+1} = (s - a_k) p_k - b_k p_{k-1} on the discrete measure */
        static __thread long double p0[SR_MAX_POINTS], p1[SR_MAX_POINTS];
        long double a[OC_MAXROOTS], b[OC_MAXROOTS];
        long double nrm_prev = 0, mu0 = 0;
        for (int m = 0; m < M; m++) {
                p0[m] = 0;
                p1[m] = 1;
                mu0 += pw[m];
        }
        for (int k = 0; k < n; k++) {
                long double nrm = 0, sa = 0;
                for (int m = 0; m < M; m++) {
                        long double q = pw[m] * p1[m] * p1[m];
                        nrm += q;
                        sa += q * ps[m];
                }
                a[k] = sa / nrm;
                b[k] = k ? nrm / nrm_prev : nrm;
                nrm_prev = nrm;
                for (int m = 0; m < M; m++) {
                        long double pn = (ps[m] - a[k]) * p1[m] - (k ? b[k] * p0[m] : 0);
                        p0[m] = p1[m];
                        p1[m] = pn;
                }
        }
        __float128 d[OC_MAXROOTS], e[OC_MAXROOTS], z[OC_MAXROOTS];
        for (int i = 0; i < n; i++) {
                d[i] = a[i];
                e[i] = (i + 1 < n) ? sqrtq((__float128)b[i + 1]) : 0;
        }
        ql_first_row(n, d, e, z);
        int idx[OC_MAXROOTS];
        for (int i = 0; i < n; i++) idx[i] = i;
        for (int i = 1; i < n; i++)
                for (int j = i; j > 0 && d[idx[j]] < d[idx[j - 1]]; j--) {
                        int t = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = t;
                }
        for (int i = 0; i < n; i++) {
                __float128 sv = d[idx[i]];
                u[i] = (double)(sv / (1 - sv));
                w[i] = (double)((__float128)mu0 * z[idx[i]] * z[idx[i]]);
        }
}

void orc_sr_rys_roots(int n, double x, double lower, double *u, double *w)
{
        rys_sr_rule(n, x, lower, u, w);
}

static void gauss_from_moments_q(int n, const __float128 *mu, __float128 *s_out, __float128 *w_out)
{
        __float128 sig_prev[2 * OC_MAXROOTS], sig[2 * OC_MAXROOTS], sig_new[2 * OC_MAXROOTS];
        __float128 alpha[OC_MAXROOTS], beta[OC_MAXROOTS];
        for (int k = 0; k < 2 * n; k++) {
                sig_prev[k] = 0;
                sig[k] = mu[k];
        }
        alpha[0] = mu[1] / mu[0];
        beta[0] = mu[0];
        for (int k = 1; k < n; k++) {
                for (int l = k; l < 2 * n - k; l++)
                        sig_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l];
                alpha[k] = sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1];
                beta[k] = sig_new[k] / sig[k - 1];
                for (int l = 0; l < 2 * n; l++) {
                        sig_prev[l] = sig[l];
                        sig[l] = sig_new[l];
                }
        }
        __float128 d[OC_MAXROOTS], e[OC_MAXROOTS], z[OC_MAXROOTS];
        for (int i = 0; i < n; i++) {
                d[i] = alpha[i];
                e[i] = (i + 1 < n) ? sqrtq(beta[i + 1]) : 0;
        }
        ql_first_row(n, d, e, z);
        /* ascending order */
        int idx[OC_MAXROOTS];
        for (int i = 0; i < n; i++) idx[i] = i;
        for (int i = 1; i < n; i++)
                for (int j = i; j > 0 && d[idx[j]] < d[idx[j - 1]]; j--) {
                        int t = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = t;
                }
        for (int i = 0; i < n; i++) {
                s_out[i] = d[idx[i]];
                w_out[i] = mu[0] * z[idx[i]] * z[idx[i]];
        }
}

void orc_rys_roots_exact(int n, double x, double *u, double *w)
{
        __float128 s[OC_MAXROOTS], wq[OC_MAXROOTS];
        rys_exact_q(n, x, s, wq);
        for (int i = 0; i < n; i++) {
                u[i] = (double)(s[i] / (1 - s[i]));
                w[i] = (double)wq[i];
        }
}

void orc_rys_roots_table(int n, double x, double *u, double *w)
{
        if (x >= RYS_XMAX(n)) {
                const double *r = rys_lag_r + rys_lag_offset[n];
                const double *ww = rys_lag_w + rys_lag_offset[n];
                double rsx = 1.0 / sqrt(x);
                for (int i = 0; i < n; i++) {
                        u[i] = r[i] / (x - r[i]);
                        w[i] = ww[i] * rsx;
                }
                return;
        }
        int it = (int)x;
        double z = 2.0 * (x - it) - 1.0;
        const double *c = rys_coef + rys_table_offset[n] + (size_t)it * RYS_NCOEF * 2 * n;
        for (int f = 0; f < 2 * n; f++) {
                double acc = c[(RYS_NCOEF - 1) * 2 * n + f];
                for (int j = RYS_NCOEF - 2; j >= 0; j--)
                        acc = acc * z + c[j * 2 * n + f];
                if (f < n) u[f] = acc;
                else w[f - n] = acc;
        }
}

int orc_rys_nmax_table(void) { return RYS_NMAX; }

static void rys_roots(int n, double x, double *u, double *w)
{
        if (g_root_mode == ORC_ROOTS_TABLE && n <= RYS_NMAX)
                orc_rys_roots_table(n, x, u, w);
        else
                orc_rys_roots_exact(n, x, u, w);
}

/* ------------------------------------------------------------------ */
/* Cartesian -> real spherical coefficients.                           */
/* Real solid harmonics (Helgaker, Jorgensen, Olsen, eq. 6.4.47-50)    */
/* scaled by sqrt((2l+1)/(4 pi)); s and p are the identity because     */
/* their angular factors sit in the common factor (src/g1e.c:565-572). */
/* Layout: c[m][cart], m = -l..l, cart in libcint order (lx desc, ly   */
/* desc); same convention as src/cart2sph.c:21-812.                   */
/* ------------------------------------------------------------------ */
static double fact(int n) { double r = 1; for (int i = 2; i <= n; i++) r *= i; return r; }
static double binom(int n, int k)
{
        if (k < 0 || k > n) return 0;
        return fact(n) / (fact(k) * fact(n - k));
}
static int cart_index(int l, int lx, int ly, int lz)
{
        (void)ly;
        int m = l - lx;
        return m * (m + 1) / 2 + lz;
}

void orc_c2s_matrix(int l, double *c)
{
        int nc = (l + 1) * (l + 2) / 2;
        int ns = 2 * l + 1;
        memset(c, 0, sizeof(double) * nc * ns);
        if (l == 0) { c[0] = 1; return; }
        if (l == 1) { c[0] = 1; c[4] = 1; c[8] = 1; return; }
        double pref = sqrt((2 * l + 1) / (4 * OC_PI));
        for (int m = -l; m <= l; m++) {
                int am = abs(m);
                double nlm = 1.0 / (pow(2, am) * fact(l)) *
                             sqrt(2 * fact(l + am) * fact(l - am) / (m == 0 ? 2.0 : 1.0));
                /* v runs over integers (m >= 0) or half-integers (m < 0):
                 * work with vv = 2v */
                int vm2 = (m < 0) ? 1 : 0;
                for (int t = 0; t <= (l - am) / 2; t++)
                for (int u = 0; u <= t; u++)
                for (int vv = vm2; vv <= am; vv += 2) {
                        double sign = (((t + (vv - vm2) / 2) & 1) ? -1.0 : 1.0);
                        double coef = sign * pow(0.25, t) * binom(l, t) * binom(l - t, am + t) *
                                      binom(t, u) * binom(am, vv);
                        int px = 2 * t + am - 2 * u - vv;
                        int py = 2 * u + vv;
                        int pz = l - 2 * t - am;
                        if (px < 0 || py < 0 || pz < 0) continue;
                        c[(m + l) * nc + cart_index(l, px, py, pz)] += pref * nlm * coef;
                }
        }
}

double orc_gto_norm(int l, double a)
{
        double n1 = l + 1.5;
        return 1.0 / sqrt(exp(lgamma(n1)) / (2.0 * pow(2 * a, n1)));
}

static double common_fac_sp(int l)
{
        /* 1/(2 sqrt(pi)) and sqrt(3/(4 pi)), src/g1e.c:565-572 */
        if (l == 0) return 0.282094791773878143;
        if (l == 1) return 0.488602511902919921;
        return 1;
}

/* ------------------------------------------------------------------ */
/* Integral environment (src/g2e.c:21-172, src/g3c2e.c:21-132)          */
/* ------------------------------------------------------------------ */
typedef struct {
        int l[4];      /* angular momenta i, j, k, l */
        int lc[4];     /* ceilings (+1 on i for the ip1 derivative) */
        int nf[4];     /* Cartesian components per shell */
        int nfall;
        int nroots;    /* roots per primitive (2 * rys_order for doubled SR) */
        int rys_order;
        int ibase, kbase;
        int dli, dlj, dlk, dll;
        int gsi, gsk, gsl, gsj; /* g strides (root fastest) */
        int gsize;
        int nmax, mmax, dn, dm;
        double rirj[3], rkrl[3];
        const double *rxij, *rxkl;
        const double *r[4];
        double common_factor;
        double expcutoff;
        double omega;
        int three_center;
        int two_center; /* 2c2e: the j slot is an exponent-0 s function too */
        int deriv_i; /* 1 for int2e_ip1 / int3c2e_ip1 */
        int deriv_k; /* 1 for int3c2e_ip2 (the auxiliary shell) or int2e_ip2 (shell k) */
        int ncomp;
} OcEnv;

static void oc_strides(OcEnv *e)
{
        int n = e->nroots;
        e->ibase = e->lc[0] > e->lc[1];
        if (e->ibase) { e->dli = e->lc[0] + e->lc[1] + 1; e->dlj = e->lc[1] + 1; }
        else { e->dli = e->lc[0] + 1; e->dlj = e->lc[0] + e->lc[1] + 1; }
        if (e->three_center) {
                e->kbase = 0;
                e->dlk = e->lc[3] + 1;
                e->dll = 1;
                e->gsi = n;
                e->gsk = n * e->dli;
                e->gsl = n * e->dli;
                e->gsj = n * e->dli * e->dlk;
                e->gsize = n * e->dli * e->dlk * e->dlj;
        } else {
                e->kbase = e->lc[2] > e->lc[3];
                if (e->kbase) { e->dlk = e->lc[2] + e->lc[3] + 1; e->dll = e->lc[3] + 1; }
                else { e->dlk = e->lc[2] + 1; e->dll = e->lc[2] + e->lc[3] + 1; }
                e->gsi = n;
                e->gsk = n * e->dli;
                e->gsl = n * e->dli * e->dlk;
                e->gsj = n * e->dli * e->dlk * e->dll;
                e->gsize = n * e->dli * e->dlk * e->dll * e->dlj;
        }
        e->nmax = e->lc[0] + e->lc[1];
        e->mmax = e->lc[2] + e->lc[3];
        e->dn = e->ibase ? e->gsi : e->gsj;
        if (e->three_center) e->dm = e->gsk;
        else e->dm = e->kbase ? e->gsk : e->gsl;
        for (int d = 0; d < 3; d++) {
                if (e->ibase) e->rirj[d] = e->r[0][d] - e->r[1][d];
                else e->rirj[d] = e->r[1][d] - e->r[0][d];
                if (e->three_center) e->rkrl[d] = 0;
                else if (e->kbase) e->rkrl[d] = e->r[2][d] - e->r[3][d];
                else e->rkrl[d] = e->r[3][d] - e->r[2][d];
        }
        e->rxij = e->ibase ? e->r[0] : e->r[1];
        if (e->three_center) e->rxkl = e->r[2];
        else e->rxkl = e->kbase ? e->r[2] : e->r[3];
}

/* Four-centre (ij|kl), src/g2e.c:21-172; deriv_i raises i for int2e_ip1
 * (ng = {1,0,0,0,...}, src/autocode/grad2.c:69-88), deriv_k raises k for
 * int2e_ip2 (ng = {0,0,1,0,...}, src/autocode/grad2.c:151-170); int2e_ipip1
 * raises i by two (ng = {2,0,0,0,...}, src/autocode/hess.c:1175-1194) and
 * int2e_ip1ip2 raises i and k by one (ng = {1,0,1,0,...}, hess.c:1339-1351). */
static void oc_init_2e(OcEnv *e, const int *shls, const int *atm, const int *bas,
                       const double *env, int deriv_i, int deriv_k)
{
        memset(e, 0, sizeof(*e));
        for (int s = 0; s < 4; s++) {
                e->l[s] = BAS(ANG_OF, shls[s]);
                e->lc[s] = e->l[s];
                e->nf[s] = (e->l[s] + 1) * (e->l[s] + 2) / 2;
                e->r[s] = env + ATM(PTR_COORD, BAS(ATOM_OF, shls[s]));
        }
        e->deriv_i = deriv_i;
        e->deriv_k = deriv_k;
        e->lc[0] += deriv_i;
        e->lc[2] += deriv_k;
        /* second derivatives (ipip1: deriv_i = 2; ip1ip2: deriv_i = deriv_k = 1)
         * have nine components */
        e->ncomp = deriv_i + deriv_k == 2 ? 9 : ((deriv_i || deriv_k) ? 3 : 1);
        e->nfall = e->nf[0] * e->nf[1] * e->nf[2] * e->nf[3];
        e->common_factor = (OC_PI * OC_PI * OC_PI) * 2 / OC_SQRTPI *
                           common_fac_sp(e->l[0]) * common_fac_sp(e->l[1]) *
                           common_fac_sp(e->l[2]) * common_fac_sp(e->l[3]);
        if (env[PTR_EXPCUTOFF] == 0) e->expcutoff = OC_EXPCUTOFF;
        else e->expcutoff = fmax(OC_MIN_EXPCUTOFF, env[PTR_EXPCUTOFF]) + 1;
        e->omega = env[PTR_RANGE_OMEGA];
        e->rys_order = (e->lc[0] + e->lc[1] + e->lc[2] + e->lc[3]) / 2 + 1;
        /* SR with rys_order <= 3 runs full - long-range on doubled roots,
         * src/g2e.c:73-79 */
        e->nroots = e->rys_order * ((e->omega < 0 && e->rys_order <= 3) ? 2 : 1);
        oc_strides(e);
}

/* Three-centre (ij|k), src/g3c2e.c:21-132; deriv_i adds the derivative
 * on i (int3c2e_ip1, ng = {1,0,0,0,...}, src/autocode/int3c2e.c:72-87),
 * deriv_k the derivative on the auxiliary shell k (int3c2e_ip2, ng =
 * {0,0,1,0,...}, :153-168).  two_center: (i|k) of int2c2e
 * (src/g2c2e.c:15-90, src/cint2c2e.c:351-374), the j slot an exponent-0 s
 * function without an angular factor. */
static void oc_init_3c2e(OcEnv *e, const int *shls, const int *atm, const int *bas,
                         const double *env, int deriv_i, int deriv_k, int two_center)
{
        memset(e, 0, sizeof(*e));
        e->three_center = 1;
        e->two_center = two_center;
        for (int s = 0; s < 3; s++) {
                if (two_center && s == 1) continue;
                e->l[s] = BAS(ANG_OF, shls[s]);
                e->nf[s] = (e->l[s] + 1) * (e->l[s] + 2) / 2;
                e->r[s] = env + ATM(PTR_COORD, BAS(ATOM_OF, shls[s]));
        }
        if (two_center) {
                e->l[1] = 0;
                e->nf[1] = 1;
                e->r[1] = e->r[0];
        }
        e->l[3] = 0;
        e->nf[3] = 1;
        e->r[3] = e->r[2];
        e->deriv_i = deriv_i;
        e->deriv_k = deriv_k;
        /* the auxiliary shell occupies the 'l' slot of the 4c machinery */
        e->lc[0] = e->l[0] + deriv_i;
        e->lc[1] = e->l[1];
        e->lc[2] = 0;
        e->lc[3] = e->l[2] + deriv_k;
        e->ncomp = (deriv_i || deriv_k) ? 3 : 1;
        e->nfall = e->nf[0] * e->nf[1] * e->nf[2];
        e->common_factor = (OC_PI * OC_PI * OC_PI) * 2 / OC_SQRTPI *
                           common_fac_sp(e->l[0]) * (two_center ? 1.0 : common_fac_sp(e->l[1])) *
                           common_fac_sp(e->l[2]);
        if (env[PTR_EXPCUTOFF] == 0) e->expcutoff = OC_EXPCUTOFF;
        else e->expcutoff = fmax(OC_MIN_EXPCUTOFF, env[PTR_EXPCUTOFF]);
        e->omega = env[PTR_RANGE_OMEGA];
        e->rys_order = (e->lc[0] + e->lc[1] + e->lc[3]) / 2 + 1;
        e->nroots = e->rys_order * ((e->omega < 0 && e->rys_order <= 3) ? 2 : 1);
        oc_strides(e);
}

/* nabla on the auxiliary shell of a three-centre g (its 'l' slot):
 * f(k) = k g(k-1) - 2 ak g(k+1), as src/g2e.c:4550-4591 does for i */
static void oc_nabla_aux(double *f, const double *g, const OcEnv *e, double ak)
{
        int nr = e->nroots;
        double ak2 = -2 * ak;
        for (int d = 0; d < 3; d++) {
                const double *gd = g + d * e->gsize;
                double *fd = f + d * e->gsize;
                for (int j = 0; j <= e->l[1]; j++)
                for (int k = 0; k <= e->l[2]; k++)
                for (int i = 0; i <= e->l[0]; i++)
                for (int r = 0; r < nr; r++) {
                        int p = r + i * e->gsi + k * e->gsl + j * e->gsj;
                        double v = ak2 * gd[p + e->gsl];
                        if (k > 0) v += k * gd[p - e->gsl];
                        fd[p] = v;
                }
        }
}

/* Cartesian exponents in libcint order: lx descending, then ly descending
 * (src/cint_bas.c:166-180). */
static void cart_comp(int l, int *nx, int *ny, int *nz)
{
        int n = 0;
        for (int lx = l; lx >= 0; lx--)
                for (int ly = l - lx; ly >= 0; ly--) {
                        nx[n] = lx; ny[n] = ly; nz[n] = l - lx - ly; n++;
                }
}

/* gout offsets, components ordered (i fastest, k, l, j) as src/g2e.c:174-266 */
static void oc_index_xyz(const OcEnv *e, int *idx)
{
        int inx[136], iny[136], inz[136], jnx[136], jny[136], jnz[136];
        int knx[136], kny[136], knz[136], lnx[136], lny[136], lnz[136];
        cart_comp(e->l[0], inx, iny, inz);
        cart_comp(e->l[1], jnx, jny, jnz);
        if (e->three_center) {
                /* aux shell lives in the l slot, k slot is an s function */
                cart_comp(0, knx, kny, knz);
                cart_comp(e->l[2], lnx, lny, lnz);
        } else {
                cart_comp(e->l[2], knx, kny, knz);
                cart_comp(e->l[3], lnx, lny, lnz);
        }
        int nfk = e->three_center ? 1 : e->nf[2];
        int nfl = e->three_center ? e->nf[2] : e->nf[3];
        int n = 0;
        for (int j = 0; j < e->nf[1]; j++)
        for (int l = 0; l < nfl; l++)
        for (int k = 0; k < nfk; k++)
        for (int i = 0; i < e->nf[0]; i++) {
                idx[n++] = e->gsj * jnx[j] + e->gsl * lnx[l] + e->gsk * knx[k] + e->gsi * inx[i];
                idx[n++] = e->gsize + e->gsj * jny[j] + e->gsl * lny[l] + e->gsk * kny[k] + e->gsi * iny[i];
                idx[n++] = 2 * e->gsize + e->gsj * jnz[j] + e->gsl * lnz[l] + e->gsk * knz[k] + e->gsi * inz[i];
        }
}

/* 2D recurrence, src/g2e.c:272-421: g(root, n, m) for n <= nmax, m <= mmax */
static void oc_vrr(double *g, const OcEnv *e, const double *b00, const double *b10,
                   const double *b01, const double *c00, const double *c0p)
{
        int nr = e->nroots, dn = e->dn, dm = e->dm;
        for (int d = 0; d < 3; d++) {
                double *gd = g + d * e->gsize;
                for (int r = 0; r < nr; r++) {
                        if (d < 2) gd[r] = 1; /* z carries the weight */
                        double cn = c00[d * nr + r], cm = c0p[d * nr + r];
                        for (int n = 0; n < e->nmax; n++) {
                                double v = cn * gd[r + n * dn];
                                if (n > 0) v += n * b10[r] * gd[r + (n - 1) * dn];
                                gd[r + (n + 1) * dn] = v;
                        }
                        for (int m = 0; m < e->mmax; m++) {
                                for (int n = 0; n <= e->nmax; n++) {
                                        double v = cm * gd[r + n * dn + m * dm];
                                        if (m > 0) v += m * b01[r] * gd[r + n * dn + (m - 1) * dm];
                                        if (n > 0) v += n * b00[r] * gd[r + (n - 1) * dn + m * dm];
                                        gd[r + n * dn + (m + 1) * dm] = v;
                                }
                        }
                }
        }
}

/* Transfer of angular momentum between the two centres of a pair,
 * src/g2e.c:428-693:  g(a+1, b) = rab * g(a, b) + g(a, b+1). */
static void oc_hrr(double *g, const OcEnv *e)
{
        int nr = e->nroots;
        for (int d = 0; d < 3; d++) {
                double *gd = g + d * e->gsize;
                /* bra: base centre is i (ibase) or j */
                int lto = e->ibase ? e->lc[1] : e->lc[0];
                int sto = e->ibase ? e->gsj : e->gsi;  /* stride of the target index */
                int sfr = e->ibase ? e->gsi : e->gsj;  /* stride of the base index */
                double rx = e->rirj[d];
                /* before the ket transfer only the ket base index is populated
                 * (0..mmax); the other ket index is 0 */
                int kmax = e->three_center ? 0 : (e->kbase ? e->mmax : 0);
                int lmax = e->three_center ? e->mmax : (e->kbase ? 0 : e->mmax);
                for (int t = 1; t <= lto; t++)
                for (int a = 0; a <= e->nmax - t; a++)
                for (int k = 0; k <= kmax; k++)
                for (int l = 0; l <= lmax; l++)
                for (int r = 0; r < nr; r++) {
                        int p = r + k * e->gsk + l * e->gsl + a * sfr + t * sto;
                        gd[p] = rx * gd[p - sto] + gd[p - sto + sfr];
                }
                if (e->three_center) continue;
                int kto = e->kbase ? e->lc[3] : e->lc[2];
                int skto = e->kbase ? e->gsl : e->gsk;
                int skfr = e->kbase ? e->gsk : e->gsl;
                int nbi = e->lc[0], nbj = e->lc[1];
                rx = e->rkrl[d];
                for (int t = 1; t <= kto; t++)
                for (int a = 0; a <= e->mmax - t; a++)
                for (int j = 0; j <= nbj; j++)
                for (int i = 0; i <= nbi; i++)
                for (int r = 0; r < nr; r++) {
                        int p = r + i * e->gsi + j * e->gsj + a * skfr + t * skto;
                        gd[p] = rx * gd[p - skto] + gd[p - skto + skfr];
                }
        }
}

/* Per primitive quartet: roots, Rys coefficients, VRR, HRR.
 * src/g2e.c:4425-4545.  Returns 0 when nothing is computed. */
static int oc_g0_2e(double *g, const OcEnv *e, double aij, double akl,
                    const double *rij, const double *rkl, double fac, double cutoff)
{
        int nr = e->nroots;
        double u[OC_MAXROOTS], w[OC_MAXROOTS];
        double xijkl[3] = {rij[0] - rkl[0], rij[1] - rkl[1], rij[2] - rkl[2]};
        double rr = xijkl[0] * xijkl[0] + xijkl[1] * xijkl[1] + xijkl[2] * xijkl[2];
        double a1 = aij * akl;
        double a0 = a1 / (aij + akl);
        double fac1 = sqrt(a0 / (a1 * a1 * a1)) * fac;
        double x = a0 * rr;
        if (e->omega == 0) {
                rys_roots(nr, x, u, w);
        } else if (e->omega > 0) {
                /* long-range erf(omega r)/r, src/g2e.c:4477-4492 */
                double theta = e->omega * e->omega / (e->omega * e->omega + a0);
                rys_roots(nr, x * theta, u, w);
                fac1 *= sqrt(theta);
                for (int r = 0; r < nr; r++) {
                        double ut = u[r] * theta;
                        u[r] = ut / (u[r] + 1. - ut);
                }
        } else {
                /* short-range erfc(|omega| r)/r, src/g2e.c:4447-4476 */
                double theta = e->omega * e->omega / (e->omega * e->omega + a0);
                if (theta * x > cutoff || theta * x > OC_EXPCUTOFF_SR) return 0;
                int ro = e->rys_order;
                if (ro == nr) {
                        orc_sr_rys_roots(nr, x, sqrt(theta), u, w);
                } else {
                        rys_roots(ro, x, u, w);
                        rys_roots(ro, theta * x, u + ro, w + ro);
                        double mst = -sqrt(theta);
                        for (int r = ro; r < nr; r++) {
                                double ut = u[r] * theta;
                                u[r] = ut / (u[r] + 1. - ut);
                                w[r] *= mst;
                        }
                }
        }
        double *gz = g + 2 * e->gsize;
        double b00[OC_MAXROOTS], b10[OC_MAXROOTS], b01[OC_MAXROOTS];
        double c00[3 * OC_MAXROOTS], c0p[3 * OC_MAXROOTS];
        double rijrx[3], rklrx[3];
        for (int d = 0; d < 3; d++) {
                rijrx[d] = rij[d] - e->rxij[d];
                rklrx[d] = rkl[d] - e->rxkl[d];
        }
        for (int r = 0; r < nr; r++) {
                double u2 = a0 * u[r];
                double tmp4 = .5 / (u2 * (aij + akl) + a1);
                double tmp5 = u2 * tmp4;
                double tmp1 = 2. * tmp5;
                double tmp2 = tmp1 * akl;
                double tmp3 = tmp1 * aij;
                b00[r] = tmp5;
                b10[r] = tmp5 + tmp4 * akl;
                b01[r] = tmp5 + tmp4 * aij;
                for (int d = 0; d < 3; d++) {
                        c00[d * nr + r] = rijrx[d] - tmp2 * xijkl[d];
                        c0p[d * nr + r] = rklrx[d] + tmp3 * xijkl[d];
                }
                gz[r] = w[r] * fac1;
        }
        oc_vrr(g, e, b00, b10, b01, c00, c0p);
        oc_hrr(g, e);
        return 1;
}

static void oc_gout_plain(double *gout, const double *g, const int *idx, const OcEnv *e)
{
        int nr = e->nroots;
        for (int n = 0; n < e->nfall; n++) {
                const double *gx = g + idx[3 * n], *gy = g + idx[3 * n + 1], *gz = g + idx[3 * n + 2];
                double s = 0;
                for (int r = 0; r < nr; r++) s += gx[r] * gy[r] * gz[r];
                gout[n] += s;
        }
}

/* nabla_i of g: f(i) = i g(i-1) - 2 ai g(i+1), src/g2e.c:4550-4591, for
 * i <= li and k <= lk (the second derivatives take the first one a step
 * further on the shell differentiated again: G2E_D_I/G2E_D_K with i_l+1,
 * src/autocode/hess.c:1134-1136,1302-1304) */
static void oc_nabla_i_to(double *f, const double *g, const OcEnv *e, double ai, int li, int lk)
{
        int nr = e->nroots;
        double ai2 = -2 * ai;
        for (int d = 0; d < 3; d++) {
                const double *gd = g + d * e->gsize;
                double *fd = f + d * e->gsize;
                for (int j = 0; j <= e->l[1]; j++)
                for (int l = 0; l <= e->l[3]; l++)
                for (int k = 0; k <= lk; k++)
                for (int i = 0; i <= li; i++)
                for (int r = 0; r < nr; r++) {
                        int p = r + i * e->gsi + k * e->gsk + l * e->gsl + j * e->gsj;
                        double v = ai2 * gd[p + e->gsi];
                        if (i > 0) v += i * gd[p - e->gsi];
                        fd[p] = v;
                }
        }
}

static void oc_nabla_i(double *f, const double *g, const OcEnv *e, double ai)
{
        oc_nabla_i_to(f, g, e, ai, e->l[0], e->l[2]);
}

/* nabla on k of a four-centre g: f(k) = k g(k-1) - 2 ak g(k+1)
 * (CINTnabla1k_2e, src/g2e.c:4651-4697; G2E_D_K of src/autocode/grad2.c:107),
 * for i <= li and k <= lk */
static void oc_nabla_k_to(double *f, const double *g, const OcEnv *e, double ak, int li, int lk)
{
        int nr = e->nroots;
        double ak2 = -2 * ak;
        for (int d = 0; d < 3; d++) {
                const double *gd = g + d * e->gsize;
                double *fd = f + d * e->gsize;
                for (int j = 0; j <= e->l[1]; j++)
                for (int l = 0; l <= e->l[3]; l++)
                for (int k = 0; k <= lk; k++)
                for (int i = 0; i <= li; i++)
                for (int r = 0; r < nr; r++) {
                        int p = r + i * e->gsi + k * e->gsk + l * e->gsl + j * e->gsj;
                        double v = ak2 * gd[p + e->gsk];
                        if (k > 0) v += k * gd[p - e->gsk];
                        fd[p] = v;
                }
        }
}

static void oc_nabla_k(double *f, const double *g, const OcEnv *e, double ak)
{
        oc_nabla_k_to(f, g, e, ak, e->l[0], e->l[2]);
}

/* Nine components of a second derivative, src/autocode/hess.c:1125-1174
 * (int2e_ipip1) and :1289-1338 (int2e_ip1ip2): component 3p+q carries the
 * first operator along axis p and the second along q; fa is the first
 * operator applied to g, fb the second applied to g, fab the first applied
 * to fb (axis p = q).  gout[9n + 3p + q]. */
static void oc_gout_hess(double *gout, const double *g, const double *fa, const double *fb,
                         const double *fab, const int *idx, const OcEnv *e)
{
        int nr = e->nroots;
        for (int n = 0; n < e->nfall; n++) {
                int ix[3] = {idx[3 * n], idx[3 * n + 1], idx[3 * n + 2]};
                for (int p = 0; p < 3; p++)
                for (int q = 0; q < 3; q++) {
                        const double *src[3];
                        for (int d = 0; d < 3; d++)
                                src[d] = d == p ? (d == q ? fab : fa) : (d == q ? fb : g);
                        double s = 0;
                        for (int r = 0; r < nr; r++)
                                s += src[0][ix[0] + r] * src[1][ix[1] + r] * src[2][ix[2] + r];
                        gout[9 * n + 3 * p + q] += s;
                }
        }
}

static void oc_gout_ip1(double *gout, const double *g, const double *f, const int *idx,
                        const OcEnv *e)
{
        int nr = e->nroots;
        for (int n = 0; n < e->nfall; n++) {
                int ix = idx[3 * n], iy = idx[3 * n + 1], iz = idx[3 * n + 2];
                double s0 = 0, s1 = 0, s2 = 0;
                for (int r = 0; r < nr; r++) {
                        s0 += f[ix + r] * g[iy + r] * g[iz + r];
                        s1 += g[ix + r] * f[iy + r] * g[iz + r];
                        s2 += g[ix + r] * g[iy + r] * f[iz + r];
                }
                gout[3 * n + 0] += s0;
                gout[3 * n + 1] += s1;
                gout[3 * n + 2] += s2;
        }
}

static double oc_dist2(const double *a, const double *b)
{
        return (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
               (a[2] - b[2]) * (a[2] - b[2]);
}

/* pair data of one shell pair, src/optimizer.c:288-342.  Writes per
 * primitive pair (ip fastest): cce, eij, rij.  Returns 1 when every pair
 * is screened out. */
static int oc_pairdata(double *cce, double *eij, double *rij,
                       const double *ai, const double *aj, const double *ci, const double *cj,
                       int npi, int npj, int nci, int ncj,
                       const double *ri, const double *rj, int lij, double expcutoff,
                       double omega)
{
        double rr = 0;
        for (int d = 0; d < 3; d++) rr += (ri[d] - rj[d]) * (ri[d] - rj[d]);
        double a_last = ai[npi - 1] + aj[npj - 1];
        double log_rr = 1.7 - 1.5 * log(a_last);
        if (lij > 0) {
                if (omega < 0) {
                        /* short range: distance term widened by theta * r_guess */
                        double theta = omega * omega / (omega * omega + a_last);
                        log_rr += lij * log(sqrt(rr) + theta * 8. + 1.);
                } else {
                        log_rr += lij * log(sqrt(rr) + 1.);
                }
        }
        int empty = 1;
        for (int jp = 0; jp < npj; jp++) {
                double mcj = 0;
                for (int c = 0; c < ncj; c++) mcj = fmax(mcj, fabs(cj[c * npj + jp]));
                for (int ip = 0; ip < npi; ip++) {
                        double mci = 0;
                        for (int c = 0; c < nci; c++) mci = fmax(mci, fabs(ci[c * npi + ip]));
                        int n = jp * npi + ip;
                        double inv = 1 / (ai[ip] + aj[jp]);
                        double e = rr * ai[ip] * aj[jp] * inv;
                        cce[n] = e - log_rr - log(mci) - log(mcj);
                        if (cce[n] < expcutoff) {
                                empty = 0;
                                double wj = aj[jp] * inv;
                                for (int d = 0; d < 3; d++) rij[3 * n + d] = ri[d] + wj * (rj[d] - ri[d]);
                                eij[n] = exp(-e);
                        } else {
                                for (int d = 0; d < 3; d++) rij[3 * n + d] = 1e18;
                                eij[n] = 0;
                        }
                }
        }
        return empty;
}

/* Cartesian -> spherical over one index of a (n0 x nc x n1) block */
static void transform_index(double *dst, const double *src, int n0, int nc, int n1,
                            const double *cmat, int ns)
{
        for (int b = 0; b < n1; b++)
        for (int m = 0; m < ns; m++)
        for (int a = 0; a < n0; a++) {
                double s = 0;
                for (int c = 0; c < nc; c++) s += cmat[m * nc + c] * src[(b * nc + c) * n0 + a];
                dst[(b * ns + m) * n0 + a] = s;
        }
}

/* gctr block [nf] in (i, k, l, j) order -> out(i, j, k, l) Fortran with
 * leading dims; sph=1 applies the real solid harmonic transform first.
 * src/cart2sph.c:5324-5376 (c2s_sph_2e1) and 4558-4674 (dcopy_iklj). */
static void oc_c2s_4(double *out, const double *gblk, const int *lv, const int *dims,
                     int sph, double *buf1, double *buf2)
{
        int nc[4], ns[4];
        for (int s = 0; s < 4; s++) {
                nc[s] = (lv[s] + 1) * (lv[s] + 2) / 2;
                ns[s] = sph ? 2 * lv[s] + 1 : nc[s];
        }
        /* current layout (i, k, l, j) with i fastest */
        int cur[4] = {nc[0], nc[2], nc[3], nc[1]};
        const double *src = gblk;
        double *dst = buf1;
        if (sph) {
                const int order[4] = {0, 2, 3, 1}; /* shell of each axis */
                double cm[4][15 * 28];
                for (int a = 0; a < 4; a++) {
                        int s = order[a];
                        orc_c2s_matrix(lv[s], cm[a]);
                }
                for (int a = 0; a < 4; a++) {
                        int n0 = 1, n1 = 1;
                        for (int b = 0; b < a; b++) n0 *= cur[b];
                        for (int b = a + 1; b < 4; b++) n1 *= cur[b];
                        transform_index(dst, src, n0, cur[a], n1, cm[a], ns[order[a]]);
                        cur[a] = ns[order[a]];
                        src = dst;
                        dst = (dst == buf1) ? buf2 : buf1;
                }
        }
        int di = cur[0], dk = cur[1], dl = cur[2], dj = cur[3];
        for (int j = 0; j < dj; j++)
        for (int l = 0; l < dl; l++)
        for (int k = 0; k < dk; k++)
        for (int i = 0; i < di; i++)
                out[i + (size_t)dims[0] * (j + (size_t)dims[1] * (k + (size_t)dims[2] * l))] =
                        src[((j * dl + l) * dk + k) * di + i];
}

/* ------------------------------------------------------------------ */
/* Driver shared by int2e_{sph,cart}, int2e_ip1_{sph,cart}, int3c2e.   */
/* src/cint2e.c:57-269 (loop), 794-870 (drv)                           */
/* ------------------------------------------------------------------ */
/* three_center: 1 for (ij|k); 2 for the two-centre (i|k) of int2c2e, whose
 * shls are (i, k).  deriv_k: derivative on the auxiliary shell (3c) or on k
 * (four centres, int2e_ip2). */
static int oc_int2e_drv(double *out, const int *dims, const int *shls_in, const int *atm,
                        const int *bas, const double *env, int deriv_i, int sph, int three_center,
                        int deriv_k)
{
        OcEnv e;
        const int two_center = three_center == 2;
        int shls[4] = {shls_in[0], shls_in[1], shls_in[2], 0};
        if (two_center) { shls[1] = shls_in[0]; shls[2] = shls_in[1]; }
        else if (!three_center) shls[3] = shls_in[3];
        int sh[4] = {shls[0], shls[1], shls[2], three_center ? shls[2] : shls[3]};
        if (three_center) oc_init_3c2e(&e, shls, atm, bas, env, deriv_i, deriv_k, two_center);
        else oc_init_2e(&e, shls, atm, bas, env, deriv_i, deriv_k);

        int nprim[4], nctr[4];
        const double *ex[4], *cf[4];
        for (int s = 0; s < 4; s++) {
                nprim[s] = BAS(NPRIM_OF, sh[s]);
                nctr[s] = BAS(NCTR_OF, sh[s]);
                ex[s] = env + BAS(PTR_EXP, sh[s]);
                cf[s] = env + BAS(PTR_COEFF, sh[s]);
        }
        static const double zero_exp = 0.0, one_coef = 1.0;
        if (three_center) {
                /* 4th function: s with exponent 0 and unit coefficient */
                nprim[3] = 1; nctr[3] = 1; ex[3] = &zero_exp; cf[3] = &one_coef;
        }
        if (two_center) {
                nprim[1] = 1; nctr[1] = 1; ex[1] = &zero_exp; cf[1] = &one_coef;
        }
        int ncomp = e.ncomp;
        int nc = nctr[0] * nctr[1] * nctr[2] * nctr[3];
        int nfall = e.nfall;
        int lv[4] = {e.l[0], e.l[1], e.l[2], three_center ? 0 : e.l[3]};
        int counts[4];
        for (int s = 0; s < 4; s++) {
                int n = sph ? 2 * lv[s] + 1 : (lv[s] + 1) * (lv[s] + 2) / 2;
                counts[s] = n * nctr[s];
        }
        if (three_center) { counts[2] = (sph ? 2 * e.l[2] + 1 : e.nf[2]) * nctr[2]; counts[3] = 1; }
        const int *dd = dims ? dims : counts;
        size_t nout = (size_t)dd[0] * dd[1] * dd[2] * (three_center ? 1 : dd[3]);

        double *gctr = calloc((size_t)nfall * ncomp * nc, sizeof(double));
        double *gout = malloc(sizeof(double) * nfall * ncomp);
        double *g = malloc(sizeof(double) * e.gsize * 3);
        double *f = malloc(sizeof(double) * e.gsize * 3);
        const int hess = ncomp == 9;
        double *fb = hess ? calloc((size_t)e.gsize * 3, sizeof(double)) : NULL;
        double *fab = hess ? calloc((size_t)e.gsize * 3, sizeof(double)) : NULL;
        int *idx = malloc(sizeof(int) * nfall * 3);
        int npij = nprim[0] * nprim[1], npkl = nprim[2] * nprim[3];
        double *cce_ij = malloc(sizeof(double) * npij * 5);
        double *eij = cce_ij + npij, *rij = eij + npij;
        double *cce_kl = malloc(sizeof(double) * npkl * 5);
        double *ekl = cce_kl + npkl, *rkl = ekl + npkl;
        int not0 = 0;

        oc_index_xyz(&e, idx);
        double expcutoff = e.expcutoff;
        int empty_ij = oc_pairdata(cce_ij, eij, rij, ex[0], ex[1], cf[0], cf[1], nprim[0], nprim[1],
                                   nctr[0], nctr[1], e.r[0], e.r[1], e.lc[0] + e.lc[1], expcutoff,
                                   e.omega);
        if (two_center) {
                /* int2c2e screens no primitive (src/cint2c2e.c:57-147) */
                empty_ij = 0;
                for (int ip = 0; ip < nprim[0]; ip++) {
                        cce_ij[ip] = -1e300;
                        eij[ip] = 1;
                        for (int d = 0; d < 3; d++) rij[3 * ip + d] = e.r[0][d];
                }
        }
        int empty_kl = 0;
        if (three_center) {
                /* single centre ket: no pair screening (src/cint3c2e.c:160-211) */
                for (int kp = 0; kp < nprim[2]; kp++) {
                        cce_kl[kp] = -1e300;
                        ekl[kp] = 1;
                        for (int d = 0; d < 3; d++) rkl[3 * kp + d] = e.r[2][d];
                }
        } else {
                empty_kl = oc_pairdata(cce_kl, ekl, rkl, ex[2], ex[3], cf[2], cf[3], nprim[2],
                                       nprim[3], nctr[2], nctr[3], e.r[2], e.r[3],
                                       e.lc[2] + e.lc[3], expcutoff, e.omega);
        }
        /* short range: wider cutoff for the polynomial prefactors
         * (ADJUST_CUTOFF, src/cint2e.c:358-380, src/cint3c2e.c:279-296) */
        if (e.omega < 0 && e.rys_order > 1) {
                double w2 = e.omega * e.omega;
                int lij = e.lc[0] + e.lc[1];
                if (lij > 0) {
                        double d = sqrt(oc_dist2(e.r[0], e.r[1]));
                        double th = w2 / (w2 + ex[0][nprim[0] - 1] + ex[1][nprim[1] - 1]);
                        expcutoff += lij * log((d + th * 8. + 1.) / (d + 1.));
                }
                if (three_center) {
                        int lk = e.lc[3];
                        if (lk > 0) {
                                double th = w2 / (w2 + ex[2][nprim[2] - 1]);
                                expcutoff += lk * log(th * 8. + 1.);
                        }
                } else {
                        int lkl = e.lc[2] + e.lc[3];
                        if (lkl > 0) {
                                double d = sqrt(oc_dist2(e.r[2], e.r[3]));
                                double th = w2 / (w2 + ex[2][nprim[2] - 1] + ex[3][nprim[3] - 1]);
                                expcutoff += lkl * log((d + th * 8. + 1.) / (d + 1.));
                        }
                }
        }
        if (!empty_ij && !empty_kl) {
                for (int lp = 0; lp < nprim[3]; lp++)
                for (int kp = 0; kp < nprim[2]; kp++) {
                        int nkl = lp * nprim[2] + kp;
                        if (cce_kl[nkl] > expcutoff) continue;
                        double eijcut = three_center ? expcutoff : expcutoff - cce_kl[nkl];
                        double akl = ex[2][kp] + ex[3][lp];
                        for (int jp = 0; jp < nprim[1]; jp++)
                        for (int ip = 0; ip < nprim[0]; ip++) {
                                int nij = jp * nprim[0] + ip;
                                if (cce_ij[nij] > eijcut) continue;
                                double aij = ex[0][ip] + ex[1][jp];
                                double fac = e.common_factor * eij[nij] * ekl[nkl];
                                int st = oc_g0_2e(g, &e, aij, akl, rij + 3 * nij, rkl + 3 * nkl, fac,
                                                  eijcut - cce_ij[nij]);
                                if (st < 0) { not0 = -1; goto done; }
                                if (!st) continue;
                                not0 = 1;
                                memset(gout, 0, sizeof(double) * nfall * ncomp);
                                if (hess) {
                                        /* ipip1: fb = D_i g to i_l+1, fab = D_i fb;
                                         * ip1ip2: fb = D_k g to i_l+1, fab = D_i fb */
                                        if (deriv_k)
                                                oc_nabla_k_to(fb, g, &e, ex[2][kp], e.l[0] + 1, e.l[2]);
                                        else
                                                oc_nabla_i_to(fb, g, &e, ex[0][ip], e.l[0] + 1, e.l[2]);
                                        oc_nabla_i_to(fab, fb, &e, ex[0][ip], e.l[0], e.l[2]);
                                        oc_nabla_i(f, g, &e, ex[0][ip]);
                                        oc_gout_hess(gout, g, f, fb, fab, idx, &e);
                                } else if (deriv_i) {
                                        oc_nabla_i(f, g, &e, ex[0][ip]);
                                        oc_gout_ip1(gout, g, f, idx, &e);
                                } else if (deriv_k) {
                                        if (three_center) oc_nabla_aux(f, g, &e, ex[2][kp]);
                                        else oc_nabla_k(f, g, &e, ex[2][kp]);
                                        oc_gout_ip1(gout, g, f, idx, &e);
                                } else {
                                        oc_gout_plain(gout, g, idx, &e);
                                }
                                /* contraction: gctr[lc][kc][jc][ic][comp][nf] */
                                for (int lc = 0; lc < nctr[3]; lc++)
                                for (int kc = 0; kc < nctr[2]; kc++)
                                for (int jc = 0; jc < nctr[1]; jc++)
                                for (int ic = 0; ic < nctr[0]; ic++) {
                                        double c = cf[0][ic * nprim[0] + ip] * cf[1][jc * nprim[1] + jp] *
                                                   cf[2][kc * nprim[2] + kp] * cf[3][lc * nprim[3] + lp];
                                        if (c == 0) continue;
                                        double *gc = gctr + (size_t)nfall * ncomp *
                                                     (((lc * nctr[2] + kc) * nctr[1] + jc) * nctr[0] + ic);
                                        for (int n = 0; n < nfall * ncomp; n++) gc[n] += c * gout[n];
                                }
                        }
                }
        }
done:;
        /* output, component-major: out + nout * comp */
        int nmax_c = 1;
        for (int s = 0; s < 4; s++) nmax_c *= (lv[s] + 1) * (lv[s] + 2) / 2;
        double *blk = malloc(sizeof(double) * nmax_c);
        double *b1 = malloc(sizeof(double) * nmax_c), *b2 = malloc(sizeof(double) * nmax_c);
        int odims[4] = {dd[0], dd[1], dd[2], three_center ? 1 : dd[3]};
        int w4[4];
        for (int s = 0; s < 4; s++) w4[s] = sph ? 2 * lv[s] + 1 : (lv[s] + 1) * (lv[s] + 2) / 2;
        for (int comp = 0; comp < ncomp; comp++) {
                double *oc = out + nout * comp;
                for (int lc = 0; lc < nctr[3]; lc++)
                for (int kc = 0; kc < nctr[2]; kc++)
                for (int jc = 0; jc < nctr[1]; jc++)
                for (int ic = 0; ic < nctr[0]; ic++) {
                        double *po = oc + ic * w4[0] + (size_t)odims[0] * (jc * w4[1] +
                                     (size_t)odims[1] * (kc * w4[2] + (size_t)odims[2] * lc * w4[3]));
                        if (not0 <= 0) {
                                for (int l = 0; l < w4[3]; l++)
                                for (int k = 0; k < w4[2]; k++)
                                for (int j = 0; j < w4[1]; j++)
                                for (int i = 0; i < w4[0]; i++)
                                        po[i + (size_t)odims[0] * (j + (size_t)odims[1] * (k + (size_t)odims[2] * l))] = 0;
                                continue;
                        }
                        const double *gc = gctr + (size_t)nfall * ncomp *
                                           (((lc * nctr[2] + kc) * nctr[1] + jc) * nctr[0] + ic);
                        for (int n = 0; n < nfall; n++) blk[n] = gc[n * ncomp + comp];
                        if (three_center) {
                                /* layout (i, k=s, l=aux, j): treat aux as the k index of the output */
                                int lv3[4] = {lv[0], lv[1], 0, lv[2]};
                                int od3[4] = {odims[0], odims[1], 1, odims[2]};
                                oc_c2s_4(po, blk, lv3, od3, sph, b1, b2);
                        } else {
                                oc_c2s_4(po, blk, lv, odims, sph, b1, b2);
                        }
                }
        }
        free(blk); free(b1); free(b2);
        free(gctr); free(gout); free(g); free(f); free(fb); free(fab); free(idx); free(cce_ij);
        free(cce_kl);
        return not0;
}

/* libcint-named entry points (src/cint2e.c:1186-1210, src/autocode/grad2.c:69-88,
 * src/cint3c2e.c:693-707).  opt/cache arguments are accepted and ignored. */
int orc_int2e_sph(double *out, const int *dims, const int *shls, const int *atm, int natm,
                  const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 0, 1, 0, 0);
}
int orc_int2e_cart(double *out, const int *dims, const int *shls, const int *atm, int natm,
                   const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 0, 0, 0, 0);
}
int orc_int2e_ip1_sph(double *out, const int *dims, const int *shls, const int *atm, int natm,
                      const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 1, 1, 0, 0);
}
int orc_int2e_ip1_cart(double *out, const int *dims, const int *shls, const int *atm, int natm,
                       const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 1, 0, 0, 0);
}
/* (ij|d/dr k l), src/autocode/grad2.c:101-170 */
int orc_int2e_ip2_sph(double *out, const int *dims, const int *shls, const int *atm, int natm,
                      const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 0, 1, 0, 1);
}
int orc_int2e_ip2_cart(double *out, const int *dims, const int *shls, const int *atm, int natm,
                       const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 0, 0, 0, 1);
}
/* (d2/dr2 i j|k l), src/autocode/hess.c:1179-1194 */
int orc_int2e_ipip1_sph(double *out, const int *dims, const int *shls, const int *atm, int natm,
                        const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 2, 1, 0, 0);
}
int orc_int2e_ipip1_cart(double *out, const int *dims, const int *shls, const int *atm, int natm,
                         const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 2, 0, 0, 0);
}
/* (d/dr i j|d/dr k l), src/autocode/hess.c:1343-1358 */
int orc_int2e_ip1ip2_sph(double *out, const int *dims, const int *shls, const int *atm, int natm,
                         const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 1, 1, 0, 1);
}
int orc_int2e_ip1ip2_cart(double *out, const int *dims, const int *shls, const int *atm, int natm,
                          const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 1, 0, 0, 1);
}
int orc_int3c2e_sph(double *out, const int *dims, const int *shls, const int *atm, int natm,
                    const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 0, 1, 1, 0);
}
int orc_int3c2e_cart(double *out, const int *dims, const int *shls, const int *atm, int natm,
                     const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 0, 0, 1, 0);
}

/* int3c2e_ip1 (src/autocode/int3c2e.c:72-87), int3c2e_ip2 (:153-168) and
 * int2c2e (src/cint2c2e.c:351-374; shls = (i, k)). */
int orc_int3c2e_ip1_sph(double *out, const int *dims, const int *shls, const int *atm, int natm,
                        const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 1, 1, 1, 0);
}
int orc_int3c2e_ip1_cart(double *out, const int *dims, const int *shls, const int *atm, int natm,
                         const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 1, 0, 1, 0);
}
int orc_int3c2e_ip2_sph(double *out, const int *dims, const int *shls, const int *atm, int natm,
                        const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 0, 1, 1, 1);
}
int orc_int3c2e_ip2_cart(double *out, const int *dims, const int *shls, const int *atm, int natm,
                         const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 0, 0, 1, 1);
}
int orc_int2c2e_sph(double *out, const int *dims, const int *shls, const int *atm, int natm,
                    const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 0, 1, 2, 0);
}
int orc_int2c2e_cart(double *out, const int *dims, const int *shls, const int *atm, int natm,
                     const int *bas, int nbas, const double *env)
{
        (void)natm; (void)nbas;
        return oc_int2e_drv(out, dims, shls, atm, bas, env, 0, 0, 2, 0);
}

/* Batch driver shaped like examples/time_c2h6.c:719-745: OpenMP over a list
 * of shell quartets, dynamic schedule.  kind: 0 sph, 1 cart, 2 ip1_sph,
 * 3 ip1_cart, 4 3c2e_sph, 5 3c2e_cart, 6/7 3c2e_ip1 sph/cart, 8/9 3c2e_ip2
 * sph/cart, 10/11 2c2e sph/cart, 12/13 int2e_ip2 sph/cart (shls stride 4; 4th ignored for 3c, the
 * first two used for 2c).  out[off[q]] receives
 * quartet q packed.  Returns the number of quartets whose status was < 0. */
int orc_batch(int kind, double *out, const int64_t *off, const int *shls, int64_t nq,
              const int *atm, int natm, const int *bas, int nbas, const double *env,
              int *not0, int nthreads)
{
        int nbad = 0;
#ifdef _OPENMP
        if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 2) reduction(+:nbad)
#endif
        for (int64_t q = 0; q < nq; q++) {
                const int *s = shls + 4 * q;
                int st;
                switch (kind) {
                case 0: st = orc_int2e_sph(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 1: st = orc_int2e_cart(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 2: st = orc_int2e_ip1_sph(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 3: st = orc_int2e_ip1_cart(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 4: st = orc_int3c2e_sph(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 5: st = orc_int3c2e_cart(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 6: st = orc_int3c2e_ip1_sph(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 7: st = orc_int3c2e_ip1_cart(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 8: st = orc_int3c2e_ip2_sph(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 9: st = orc_int3c2e_ip2_cart(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 10: st = orc_int2c2e_sph(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 12: st = orc_int2e_ip2_sph(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 13: st = orc_int2e_ip2_cart(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 14: st = orc_int2e_ipip1_sph(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 15: st = orc_int2e_ipip1_cart(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 16: st = orc_int2e_ip1ip2_sph(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                case 17: st = orc_int2e_ip1ip2_cart(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                default: st = orc_int2c2e_cart(out + off[q], NULL, s, atm, natm, bas, nbas, env); break;
                }
                if (not0) not0[q] = st;
                if (st < 0) nbad++;
        }
        return nbad;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
        return omp_get_max_threads();
#else
        return 1;
#endif
}

/* Nuclear gradient of E = 1/2 sum D_ij D_kl (ij|kl) - 1/4 sum D_ik D_jl (ij|kl)
 * restated on the CPU for the tests (libcint leaves this contraction to its
 * caller; the ip1 blocks are this file's int2e_ip1_sph,
 * src/autocode/grad2.c:19-88): for every (i, j, k >= l)
 *   grad[atom(i)][x] += f_kl sum_abcd (d/dr_x a b|c d)
 *                       [-2 D_ab D_cd + (D_ac D_bd + D_ad D_bc)/2],
 * f_kl = 2 for k != l, 1 otherwise.  dm: nao x nao spherical density,
 * row-major.  OpenMP over (i, j), per-thread accumulators. */
int orc_grad_ip1(const int *atm, int natm, const int *bas, int nbas, const double *env,
                 const double *dm, int nao, double *grad, int nthreads)
{
        int *ao = malloc(sizeof(int) * (nbas + 1));
        ao[0] = 0;
        for (int s = 0; s < nbas; s++)
                ao[s + 1] = ao[s] + (2 * BAS(ANG_OF, s) + 1) * BAS(NCTR_OF, s);
        if (ao[nbas] != nao) {
                free(ao);
                return -1;
        }
        for (int i = 0; i < 3 * natm; i++) grad[i] = 0;
#ifdef _OPENMP
        if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
        {
                double *g = calloc(3 * natm, sizeof(double));
                int dmax = 0;
                for (int s = 0; s < nbas; s++)
                        if (ao[s + 1] - ao[s] > dmax) dmax = ao[s + 1] - ao[s];
                double *blk = malloc(sizeof(double) * 3 * dmax * dmax * dmax * dmax);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
                for (int64_t ij = 0; ij < (int64_t)nbas * nbas; ij++) {
                        const int i = (int)(ij / nbas), j = (int)(ij % nbas);
                        for (int k = 0; k < nbas; k++)
                                for (int l = 0; l <= k; l++) {
                                        const int sh[4] = {i, j, k, l};
                                        if (!orc_int2e_ip1_sph(blk, NULL, sh, atm, natm, bas, nbas, env))
                                                continue;
                                        const int di = ao[i + 1] - ao[i], dj = ao[j + 1] - ao[j];
                                        const int dk = ao[k + 1] - ao[k], dl = ao[l + 1] - ao[l];
                                        const int64_t nb = (int64_t)di * dj * dk * dl;
                                        const double fkl = k != l ? 2.0 : 1.0;
                                        double s[3] = {0, 0, 0};
                                        for (int d = 0; d < dl; d++)
                                        for (int c = 0; c < dk; c++)
                                        for (int b = 0; b < dj; b++)
                                        for (int a = 0; a < di; a++) {
                                                const int P = ao[i] + a, Q = ao[j] + b;
                                                const int R = ao[k] + c, S = ao[l] + d;
                                                const double w = -2.0 * dm[(int64_t)P * nao + Q] * dm[(int64_t)R * nao + S]
                                                        + 0.5 * (dm[(int64_t)P * nao + R] * dm[(int64_t)Q * nao + S]
                                                                 + dm[(int64_t)P * nao + S] * dm[(int64_t)Q * nao + R]);
                                                const int64_t e = a + (int64_t)di * (b + (int64_t)dj * (c + (int64_t)dk * d));
                                                s[0] += blk[e] * w;
                                                s[1] += blk[nb + e] * w;
                                                s[2] += blk[2 * nb + e] * w;
                                        }
                                        const int at = BAS(ATOM_OF, i);
                                        for (int x = 0; x < 3; x++) g[3 * at + x] += fkl * s[x];
                                }
                }
#ifdef _OPENMP
#pragma omp critical
#endif
                for (int x = 0; x < 3 * natm; x++) grad[x] += g[x];
                free(blk);
                free(g);
        }
        free(ao);
        return 0;
}
