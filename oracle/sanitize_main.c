/* Sanitizer driver for the CPU oracle (test infrastructure).
 *
 * Built together with cint_oracle.c under AddressSanitizer and
 * UndefinedBehaviorSanitizer (make -C oracle sanitize): reads a basis
 * (atm/bas/env as text: natm nbas nenv, then the integers and doubles) and
 * evaluates every shell tuple of libcint's test-suite loops
 * (testsuite/test_cint.py:235-256 for the four-centre kinds, test_3c2e.py's
 * (i, j, k) and (i, k) loops over the first four shells for the three- and
 * two-centre kinds), in exact and tabulated root modes, printing one line
 * "name sum|x|" per kind.  Any sanitizer finding aborts with a nonzero exit
 * status; the caller compares the sums with libcint's golden values.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

typedef int (*intor_fn)(double *, const int *, const int *, const int *, int, const int *, int,
                        const double *);
int orc_int2e_sph(double *, const int *, const int *, const int *, int, const int *, int,
                  const double *);
int orc_int2e_cart(double *, const int *, const int *, const int *, int, const int *, int,
                   const double *);
int orc_int2e_ip1_sph(double *, const int *, const int *, const int *, int, const int *, int,
                      const double *);
int orc_int2e_ip2_sph(double *, const int *, const int *, const int *, int, const int *, int,
                      const double *);
int orc_int2e_ipip1_sph(double *, const int *, const int *, const int *, int, const int *, int,
                        const double *);
int orc_int2e_ip1ip2_sph(double *, const int *, const int *, const int *, int, const int *, int,
                         const double *);
int orc_int3c2e_sph(double *, const int *, const int *, const int *, int, const int *, int,
                    const double *);
int orc_int3c2e_ip1_sph(double *, const int *, const int *, const int *, int, const int *, int,
                        const double *);
int orc_int3c2e_ip2_sph(double *, const int *, const int *, const int *, int, const int *, int,
                        const double *);
int orc_int2c2e_sph(double *, const int *, const int *, const int *, int, const int *, int,
                    const double *);
void orc_set_root_mode(int mode);

static int nsph_tot(const int *bas, int s) { return (2 * bas[8 * s + 1] + 1) * bas[8 * s + 3]; }
static int ncart_tot(const int *bas, int s)
{
        const int l = bas[8 * s + 1];
        return (l + 1) * (l + 2) / 2 * bas[8 * s + 3];
}

int main(int argc, char **argv)
{
        if (argc != 2) {
                fprintf(stderr, "usage: %s basis.txt\n", argv[0]);
                return 2;
        }
        FILE *fh = fopen(argv[1], "r");
        if (!fh) return 2;
        int natm, nbas, nenv;
        if (fscanf(fh, "%d %d %d", &natm, &nbas, &nenv) != 3) return 2;
        int *atm = malloc(sizeof(int) * 6 * natm);
        int *bas = malloc(sizeof(int) * 8 * nbas);
        double *env = malloc(sizeof(double) * nenv);
        for (int i = 0; i < 6 * natm; i++)
                if (fscanf(fh, "%d", atm + i) != 1) return 2;
        for (int i = 0; i < 8 * nbas; i++)
                if (fscanf(fh, "%d", bas + i) != 1) return 2;
        for (int i = 0; i < nenv; i++)
                if (fscanf(fh, "%lf", env + i) != 1) return 2;
        fclose(fh);
        const int nb4 = nbas < 8 ? nbas : 8;
        struct { const char *name; intor_fn f; int axes, ncomp, sph; } kinds[] = {
                {"int2e_sph", orc_int2e_sph, 4, 1, 1},
                {"int2e_cart", orc_int2e_cart, 4, 1, 0},
                {"int2e_ip1_sph", orc_int2e_ip1_sph, 4, 3, 1},
                {"int2e_ip2_sph", orc_int2e_ip2_sph, 4, 3, 1},
                {"int2e_ipip1_sph", orc_int2e_ipip1_sph, 4, 9, 1},
                {"int2e_ip1ip2_sph", orc_int2e_ip1ip2_sph, 4, 9, 1},
                {"int3c2e_sph", orc_int3c2e_sph, 3, 1, 1},
                {"int3c2e_ip1_sph", orc_int3c2e_ip1_sph, 3, 3, 1},
                {"int3c2e_ip2_sph", orc_int3c2e_ip2_sph, 3, 3, 1},
                {"int2c2e_sph", orc_int2c2e_sph, 2, 1, 1},
        };
        for (int mode = 0; mode < 2; mode++) {
                orc_set_root_mode(mode);
                for (size_t kk = 0; kk < sizeof kinds / sizeof kinds[0]; kk++) {
                        double sum = 0;
                        int s[4] = {0, 0, 0, 0};
                        const int n3 = kinds[kk].axes == 4 ? nb4 : 4;
                        for (int l = 0; l < (kinds[kk].axes == 4 ? n3 : 1); l++)
                        for (int k = 0; k < (kinds[kk].axes == 4 ? l + 1 : (kinds[kk].axes == 3 ? n3 : 1)); k++)
                        for (int j = 0; j < n3; j++)
                        for (int i = 0; i < (kinds[kk].axes == 4 ? j + 1 : n3); i++) {
                                if (kinds[kk].axes == 4) {
                                        s[0] = i; s[1] = j; s[2] = k; s[3] = l;
                                } else if (kinds[kk].axes == 3) {
                                        s[0] = i; s[1] = j; s[2] = k; s[3] = 0;
                                } else {
                                        s[0] = i; s[1] = j; s[2] = 0; s[3] = 0;
                                }
                                size_t n = kinds[kk].ncomp;
                                for (int a = 0; a < kinds[kk].axes; a++)
                                        n *= kinds[kk].sph ? nsph_tot(bas, s[a]) : ncart_tot(bas, s[a]);
                                double *out = malloc(sizeof(double) * n);
                                kinds[kk].f(out, NULL, s, atm, natm, bas, nbas, env);
                                for (size_t q = 0; q < n; q++) sum += fabs(out[q]);
                                free(out);
                        }
                        printf("%s %s %.15g\n", mode ? "table" : "exact", kinds[kk].name, sum);
                }
        }
        free(atm);
        free(bas);
        free(env);
        return 0;
}
