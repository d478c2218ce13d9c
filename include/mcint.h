/*
 * libmcint: libcint's two-electron Rys-quadrature integrals on one AMD
 * MI355X (gfx950).  C ABI.
 *
 * The data model is libcint's (include/cint.h.in:26-67 of libcint 6.1.1):
 * atm[natm*6], bas[nbas*8] (FINT = int), env[] doubles with user data from
 * slot 20, contraction coefficients already multiplied by CINTgto_norm.
 * Outputs are column-major (di,dj,dk,dl[,comp]) blocks exactly as libcint
 * writes them.
 *
 * Three groups of entry points:
 *   1. libcint-named per-call functions (drop-in; each call runs on the GPU
 *      and synchronises, so they are correct but launch-bound);
 *   2. *_batch functions taking a list of shell tuples on the host: the
 *      drop-in for a loop over the per-call functions;
 *   3. mcint_* functions with device-resident outputs, reusable plans and
 *      HIP streams: the throughput path.
 * All functions return status codes instead of calling exit().
 */
#ifndef MCINT_H
#define MCINT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef FINT
#define FINT int
#endif
#ifndef CACHE_SIZE_T
#define CACHE_SIZE_T FINT
#endif

#if !defined HAVE_DEFINED_CINTOPT_H
#define HAVE_DEFINED_CINTOPT_H
/* Same public layout as libcint's PairData / CINTOpt (include/cint.h.in:140-152).
 * libmcint allocates a larger private object behind this header; device
 * data hangs off it.  Fields other than nbas are left NULL. */
typedef struct {
        double rij[3];
        double eij;
        double cceij;
} PairData;
typedef struct {
        FINT **index_xyz_array;
        FINT **non0ctr;
        FINT **sortedidx;
        FINT nbas;
        double **log_max_coeff;
        PairData **pairdata;
} CINTOpt;
#endif

/* Status codes */
#define MCINT_OK 0
#define MCINT_ERR_ARG (-1)       /* invalid argument (shell index, l, null pointer) */
#define MCINT_ERR_UNSUPPORTED (-2) /* class or operator outside the device tables */
#define MCINT_ERR_HIP (-3)       /* a HIP runtime call failed */
#define MCINT_ERR_NOMEM (-4)     /* device or host allocation failed */

/* Integral kinds for the plan API */
#define MCINT_INT2E_SPH 0
#define MCINT_INT2E_CART 1
#define MCINT_INT2E_IP1_SPH 2
#define MCINT_INT2E_IP1_CART 3
#define MCINT_INT3C2E_SPH 4
#define MCINT_INT3C2E_CART 5
#define MCINT_INT3C2E_IP1_SPH 6
#define MCINT_INT3C2E_IP1_CART 7
#define MCINT_INT3C2E_IP2_SPH 8
#define MCINT_INT3C2E_IP2_CART 9
#define MCINT_INT2C2E_SPH 10
#define MCINT_INT2C2E_CART 11
#define MCINT_INT2E_IP2_SPH 12
#define MCINT_INT2E_IP2_CART 13
#define MCINT_INT2E_IPIP1_SPH 14
#define MCINT_INT2E_IPIP1_CART 15
#define MCINT_INT2E_IP1IP2_SPH 16
#define MCINT_INT2E_IP1IP2_CART 17

/* ---- 1. libcint-named per-call entry points --------------------------
 * Signature of libcint's CINTIntegralFunction (include/cint_funcs.h:10-17).
 * Return: nonzero iff any primitive survived screening; the block is
 * zero-filled otherwise (src/cint2e.c:857-869).  out == NULL returns a
 * positive size hint (as libcint returns the cache size, src/cint2e.c:801-816).
 * On a device error the functions return 0 and mcint_last_error() says why. */
CACHE_SIZE_T int2e_sph(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                       FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
                       /* replaces src/cint2e.c:1186-1193 */
CACHE_SIZE_T int2e_cart(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                        FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
                        /* replaces src/cint2e.c:1202-1210 */
CACHE_SIZE_T int2e_ip1_sph(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                           FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
                           /* replaces src/autocode/grad2.c:81-88 */
CACHE_SIZE_T int2e_ip1_cart(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                            FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
                            /* replaces src/autocode/grad2.c:73-80 */
CACHE_SIZE_T int3c2e_sph(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                         FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
                         /* replaces src/cint3c2e.c:693-701 */
CACHE_SIZE_T int3c2e_cart(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                          FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
                          /* replaces src/cint3c2e.c:709-717 */
/* (i j|d/dr k l), 3 components, derivative on the third shell: out
 * (di,dj,dk,dl,3); replaces src/autocode/grad2.c:155-170 (cart 155-162,
 * sph 163-170).  Runs on the i-derivative kernels as (d/dr k l|i j). */
CACHE_SIZE_T int2e_ip2_sph(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                           FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
CACHE_SIZE_T int2e_ip2_cart(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                            FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
/* (d2/dr2 i j|k l), 9 components (component 3p+q differentiates i along
 * axes p and q): out (di,dj,dk,dl,9); replaces src/autocode/hess.c:1179-1194
 * (cart 1179-1186, sph 1187-1194). */
CACHE_SIZE_T int2e_ipip1_sph(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                             FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
CACHE_SIZE_T int2e_ipip1_cart(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                              FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
/* (d/dr i j|d/dr k l), 9 components (component 3p+q: i along axis p, k
 * along axis q): out (di,dj,dk,dl,9); replaces src/autocode/hess.c:1343-1358
 * (cart 1343-1350, sph 1351-1358). */
CACHE_SIZE_T int2e_ip1ip2_sph(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                              FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
CACHE_SIZE_T int2e_ip1ip2_cart(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                               FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
/* (d/dr i j|k), 3 components, derivative on the first shell: out
 * (di,dj,dk,3); replaces src/autocode/int3c2e.c:72-87 (cart 72-79, sph 80-87). */
CACHE_SIZE_T int3c2e_ip1_sph(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                             FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
CACHE_SIZE_T int3c2e_ip1_cart(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                              FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
/* (i j|d/dr k), derivative on the auxiliary shell: out (di,dj,dk,3);
 * replaces src/autocode/int3c2e.c:153-168 (cart 153-160, sph 161-168). */
CACHE_SIZE_T int3c2e_ip2_sph(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                             FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
CACHE_SIZE_T int3c2e_ip2_cart(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                              FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
/* Two-centre Coulomb (i|k), shls = {i, k}: out (di,dk); no primitive is
 * screened.  Replaces src/cint2c2e.c:351-389 (int2c2e_sph, _cart). */
CACHE_SIZE_T int2c2e_sph(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                         FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);
CACHE_SIZE_T int2c2e_cart(double *out, FINT *dims, FINT *shls, FINT *atm, FINT natm,
                          FINT *bas, FINT nbas, double *env, CINTOpt *opt, double *cache);

/* Optimizers (src/cint2e.c:1195-1200, src/autocode/grad2.c:69-72,
 * src/cint3c2e.c:703-707) and their destructors (src/optimizer.c:40-72). */
void int2e_optimizer(CINTOpt **opt, FINT *atm, FINT natm, FINT *bas, FINT nbas, double *env);
void int2e_ip1_optimizer(CINTOpt **opt, FINT *atm, FINT natm, FINT *bas, FINT nbas, double *env);
/* src/autocode/grad2.c:151-154 */
void int2e_ip2_optimizer(CINTOpt **opt, FINT *atm, FINT natm, FINT *bas, FINT nbas, double *env);
/* src/autocode/hess.c:1175-1178, 1339-1342 */
void int2e_ipip1_optimizer(CINTOpt **opt, FINT *atm, FINT natm, FINT *bas, FINT nbas, double *env);
void int2e_ip1ip2_optimizer(CINTOpt **opt, FINT *atm, FINT natm, FINT *bas, FINT nbas, double *env);
void int3c2e_optimizer(CINTOpt **opt, FINT *atm, FINT natm, FINT *bas, FINT nbas, double *env);
/* src/autocode/int3c2e.c:68-71, 149-152; src/cint2c2e.c:360-366 */
void int3c2e_ip1_optimizer(CINTOpt **opt, FINT *atm, FINT natm, FINT *bas, FINT nbas, double *env);
void int3c2e_ip2_optimizer(CINTOpt **opt, FINT *atm, FINT natm, FINT *bas, FINT nbas, double *env);
void int2c2e_optimizer(CINTOpt **opt, FINT *atm, FINT natm, FINT *bas, FINT nbas, double *env);
void CINTdel_optimizer(CINTOpt **opt);
void CINTdel_2e_optimizer(CINTOpt **opt);
/* An optimizer built by the functions above uploads atm/bas/env and the
 * primitive-pair data of every shell pair to the current device once
 * (src/optimizer.c:183-191,344-425 builds the CPU counterpart); per-call
 * functions given that optimizer and the same atm/bas/env then only launch
 * the tuple's kernel and copy its block back.  Given other data, or an
 * optimizer without device state, they take the self-contained path
 * (upload, plan, launch).  Thread safe: concurrent callers sharing one
 * optimizer each get their own stream and buffers from a pool.
 * mcint_optimizer_on_device (no libcint counterpart) returns 1 when opt
 * holds device state, 0 otherwise (no device, or more than 8192 primitives
 * in the basis). */
int mcint_optimizer_on_device(const CINTOpt *opt);

/* Legacy names (src/misc.h:34-83, src/c2f.h:6-64) */
FINT cint2e_sph(double *out, FINT *shls, FINT *atm, FINT natm, FINT *bas, FINT nbas,
                double *env, CINTOpt *opt);
FINT cint2e_cart(double *out, FINT *shls, FINT *atm, FINT natm, FINT *bas, FINT nbas,
                 double *env, CINTOpt *opt);
FINT cint2e_ip1_sph(double *out, FINT *shls, FINT *atm, FINT natm, FINT *bas, FINT nbas,
                    double *env, CINTOpt *opt);
void cint2e_sph_optimizer(CINTOpt **opt, FINT *atm, FINT natm, FINT *bas, FINT nbas, double *env);

/* Basis helpers (src/misc.c:72-93, src/cint_bas.c:36-160) */
double CINTgto_norm(FINT n, double a);
FINT CINTcgto_spheric(const FINT bas_id, const FINT *bas);
FINT CINTcgto_cart(const FINT bas_id, const FINT *bas);
FINT CINTtot_cgto_spheric(const FINT *bas, const FINT nbas);
FINT CINTtot_cgto_cart(const FINT *bas, const FINT nbas);
void CINTshells_spheric_offset(FINT ao_loc[], const FINT *bas, const FINT nbas);
void CINTshells_cart_offset(FINT ao_loc[], const FINT *bas, const FINT nbas);

/* Rys roots u = t^2/(1-t^2) and weights of exp(-x t^2) on [0,1], evaluated
 * on the device (replaces src/rys_roots.c:57-122, declared at
 * src/rys_roots.h:3).  nroots <= 9.  On an invalid argument or a device
 * error u and w are left untouched and mcint_last_error() says why
 * (libcint calls exit() instead). */
void CINTrys_roots(int nroots, double x, double *u, double *w);
int mcint_rys_roots_batch(int nroots, const double *x, size_t n, double *u, double *w);
/* Short-range rule: roots u = t^2/(1-t^2) and weights of exp(-x t^2) on
 * [lower, 1], evaluated on the device (replaces src/rys_roots.c:149-267,
 * declared at src/rys_roots.h:4); nroots <= 9, 0 <= lower < 1.  Built by
 * the discretised Stieltjes procedure on 1024 Gauss-Legendre samples of the
 * weight and Sturm bisection of the Jacobi matrix.  The batch form takes
 * host arrays of n (x, lower) pairs and fills n * nroots roots and weights.
 * Errors as for CINTrys_roots. */
void CINTsr_rys_roots(int nroots, double x, double lower, double *u, double *w);
int mcint_sr_rys_roots_batch(int nroots, const double *x, const double *lower, size_t n,
                             double *u, double *w);

/* ---- 2. host-pointer batches -------------------------------------------
 * out + out_offsets[q] receives quartet q, byte-identical in layout to
 * int2e_sph(out + out_offsets[q], NULL, shls + 4*q, ...).  out_offsets ==
 * NULL packs the blocks in list order.  not0 (nullable) receives each
 * quartet's return value.  hip_stream may be NULL.  Returns MCINT_OK or an
 * error code.  For int3c2e the list stride is still 4 (the 4th entry is
 * ignored). */
int int2e_sph_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                    const FINT *atm, FINT natm, const FINT *bas, FINT nbas, const double *env,
                    const CINTOpt *opt, FINT *not0, void *hip_stream);
int int2e_cart_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                     const FINT *atm, FINT natm, const FINT *bas, FINT nbas, const double *env,
                     const CINTOpt *opt, FINT *not0, void *hip_stream);
int int2e_ip1_sph_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                        const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                        const double *env, const CINTOpt *opt, FINT *not0, void *hip_stream);
int int3c2e_sph_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                      const FINT *atm, FINT natm, const FINT *bas, FINT nbas, const double *env,
                      const CINTOpt *opt, FINT *not0, void *hip_stream);
/* Cartesian twins of the two above (loops over int2e_ip1_cart,
 * src/autocode/grad2.c:73-80, and int3c2e_cart, src/cint3c2e.c:709-716). */
int int2e_ip1_cart_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                         const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                         const double *env, const CINTOpt *opt, FINT *not0, void *hip_stream);
int int3c2e_cart_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                       const FINT *atm, FINT natm, const FINT *bas, FINT nbas, const double *env,
                       const CINTOpt *opt, FINT *not0, void *hip_stream);
/* Batches of the derivative three-centre and the two-centre integrals
 * (loops over int3c2e_ip1_*, int3c2e_ip2_*, src/autocode/int3c2e.c:72-87,
 * 153-168, and int2c2e_*, src/cint2c2e.c:351-374; for int2c2e the first two
 * entries of each 4-int tuple are (i, k)). */
int int3c2e_ip1_sph_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                          const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                          const double *env, const CINTOpt *opt, FINT *not0, void *hip_stream);
int int3c2e_ip1_cart_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                           const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                           const double *env, const CINTOpt *opt, FINT *not0, void *hip_stream);
int int3c2e_ip2_sph_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                          const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                          const double *env, const CINTOpt *opt, FINT *not0, void *hip_stream);
int int3c2e_ip2_cart_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                           const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                           const double *env, const CINTOpt *opt, FINT *not0, void *hip_stream);
int int2c2e_sph_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                      const FINT *atm, FINT natm, const FINT *bas, FINT nbas, const double *env,
                      const CINTOpt *opt, FINT *not0, void *hip_stream);
int int2c2e_cart_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                       const FINT *atm, FINT natm, const FINT *bas, FINT nbas, const double *env,
                       const CINTOpt *opt, FINT *not0, void *hip_stream);
/* int2e_ip2 over a list of quartets (blocks (di,dj,dk,dl,3)) */
int int2e_ip2_sph_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                        const FINT *atm, FINT natm, const FINT *bas, FINT nbas, const double *env,
                        const CINTOpt *opt, FINT *not0, void *hip_stream);
int int2e_ip2_cart_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                         const FINT *atm, FINT natm, const FINT *bas, FINT nbas, const double *env,
                         const CINTOpt *opt, FINT *not0, void *hip_stream);
/* int2e_ipip1 and int2e_ip1ip2 over a list of quartets (blocks (di,dj,dk,dl,9)) */
int int2e_ipip1_sph_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                          const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                          const double *env, const CINTOpt *opt, FINT *not0, void *hip_stream);
int int2e_ipip1_cart_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                           const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                           const double *env, const CINTOpt *opt, FINT *not0, void *hip_stream);
int int2e_ip1ip2_sph_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                           const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                           const double *env, const CINTOpt *opt, FINT *not0, void *hip_stream);
int int2e_ip1ip2_cart_batch(double *out, const size_t *out_offsets, const FINT *shls, size_t nq,
                            const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                            const double *env, const CINTOpt *opt, FINT *not0, void *hip_stream);

/* ---- 3. device-resident plans ------------------------------------------- */
typedef struct mcint_basis mcint_basis;
typedef struct mcint_plan mcint_plan;

/* Upload atm/bas/env (host pointers) to the current device. */
int mcint_basis_create(mcint_basis **b, const FINT *atm, FINT natm, const FINT *bas, FINT nbas,
                       const double *env);
void mcint_basis_destroy(mcint_basis *b);

/* Build a plan for a host list of shell tuples (stride 4): canonical class
 * buckets, primitive-pair data, launch table.  out_offsets (host, nullable
 * = packed) in doubles; a negative offset returns MCINT_ERR_ARG.
 * *out_size receives the number of doubles the output must hold. */
int mcint_plan_create(mcint_plan **p, mcint_basis *b, int intor, const FINT *shls, size_t nq,
                      const int64_t *out_offsets, int64_t *out_size);
/* Mixed precision: as mcint_plan_create, plus fp32_flags[q] != 0 marks
 * quartets whose recurrences and primitive contraction may run in FP32
 * (roots, weights, pair factors and the stored result stay FP64).  Only
 * classes with a single-chunk register kernel honour the flag; others stay
 * FP64.  mcint_plan_fp32_tasks() returns how many tasks run in FP32. */
int mcint_plan_create_mixed(mcint_plan **p, mcint_basis *b, int intor, const FINT *shls, size_t nq,
                            const int64_t *out_offsets, const uint8_t *fp32_flags,
                            int64_t *out_size);
int64_t mcint_plan_fp32_tasks(const mcint_plan *p);

/* Run the plan: out_dev and not0_dev (nullable) are device pointers. The
 * first call on given (out, not0, stream) captures a HIP graph that later
 * calls replay.  Asynchronous on hip_stream. */
int mcint_plan_execute(mcint_plan *p, double *out_dev, FINT *not0_dev, void *hip_stream);
/* Same, launching the kernels directly (no graph). */
int mcint_plan_execute_direct(mcint_plan *p, double *out_dev, FINT *not0_dev, void *hip_stream);
/* No counterpart in libcint.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops): as
 * mcint_plan_execute_direct, but the chunked and generic paths' per-task
 * blocks and flags live in a fixed device array of
 * MCINT_NOALLOC_SCRATCH_DOUBLES doubles (64 MiB per GPU, held until the
 * process ends) instead of the plan's lazily allocated scratch; buckets run
 * in slices that fit it.  Calls no hipMalloc/hipFree.  Returns
 * MCINT_ERR_UNSUPPORTED, before any launch, when one task of some class
 * needs more block space than the array holds.  Calls that overlap on one
 * GPU share the array and race. */
#define MCINT_NOALLOC_SCRATCH_DOUBLES (8ll << 20)
int mcint_plan_execute_noalloc(mcint_plan *p, double *out_dev, FINT *not0_dev, void *hip_stream);
int64_t mcint_noalloc_scratch_doubles(void);
void mcint_plan_destroy(mcint_plan *p);

/* Plan statistics: n[0] tasks, n[1] buckets, n[2] tasks on fast kernels,
 * n[3] tasks on the generic kernel, n[4] primitive quartets before
 * screening, n[5] output doubles. */
int mcint_plan_stats(const mcint_plan *p, int64_t *n);

/* Buckets of a plan (one kernel launch each).  info receives: la, lb, lc,
 * ld (canonical class), perm, team lanes, kernel path (0 generic, 1 single
 * register kernel, 2 chunked), task count. */
int mcint_plan_nbuckets(const mcint_plan *p);
int mcint_plan_bucket_info(const mcint_plan *p, int bucket, int64_t *info);
int mcint_plan_execute_bucket(mcint_plan *p, int bucket, double *out_dev, FINT *not0_dev,
                              void *hip_stream);
/* List positions (indices into the plan's shls) of tasks first..first+n-1
 * of a bucket, in launch order (host array of n ints; synchronous).  A
 * generally contracted quartet appears once per contraction column. */
int mcint_plan_bucket_tasks(const mcint_plan *p, int bucket, int64_t first, int64_t n,
                            int *quartet);
/* Primitive quartets that survive libcint's screening, per bucket (host
 * array of nbuckets counts).  Synchronous. */
int mcint_plan_count_primitives(mcint_plan *p, int64_t *per_bucket);

/* Nuclear gradient digest of an int2e_ip1_sph plan whose list holds every
 * (i, j, k>=l) shell tuple: adds to grad_dev[3*atom+x] (device, natm*3)
 *   dE/dA_x = sum_{i on A} sum_{jkl} (d/dr_x i j|kl) [-2 D_ij D_kl + D_ik D_jl]
 * for E = 1/2 sum D_ij D_kl (ij|kl) - 1/4 sum D_ik D_jl (ij|kl), without
 * storing the derivative integrals.  dm_dev: nao x nao spherical density
 * (row-major, symmetric).  Asynchronous on hip_stream. */
int mcint_plan_execute_grad(mcint_plan *p, const double *dm_dev, int nao, double *grad_dev,
                            void *hip_stream);

/* Coulomb and exchange matrices fused into the integral kernels: executes
 * an int2e_sph plan over unique quartets (i>=j, k>=l, ij>=kl, each once)
 * and, instead of storing the blocks, contracts each block with the
 * density as it is formed:
 *   vj[p,q] = sum_rs (pq|rs) D[r,s],   vk[p,r] = sum_qs (pq|rs) D[q,s]
 * (overwritten; either nullable).  dm_dev: nao x nao spherical density
 * (row-major, symmetric).  No ERI buffer: device memory is the plan's
 * scratch and two nao x nao matrices.  Plain and long-range operators;
 * mixed-precision plans digest their FP32 blocks too.  Asynchronous on
 * hip_stream.  (The integral-side layer libcint leaves to its callers,
 * PySCF's _vhf.) */
int mcint_plan_execute_jk(mcint_plan *p, const double *dm_dev, int nao, double *vj_dev,
                          double *vk_dev, void *hip_stream);

/* Coulomb / exchange matrices from computed int2e_sph blocks: for every
 * quartet q of shls_dev (4 ints each, device) with its block at
 * eri_dev + off_dev[q], and every distinct permutation (pq|rs) of it,
 *   vj[p,q] += (pq|rs) D[r,s],   vk[p,r] += (pq|rs) D[q,s].
 * Pass the unique quartets (i>=j, k>=l, ij>=kl) once each to get the full
 * J = sum_rs (pq|rs) D_rs and K = sum_qs (pq|rs) D_qs.  vj_dev / vk_dev
 * (nao x nao, row-major, nullable) are accumulated into, not zeroed. */
int mcint_jk_digest(const FINT *shls_dev, const int64_t *off_dev, size_t nq, const double *eri_dev,
                    const mcint_basis *b, const double *dm_dev, double *vj_dev, double *vk_dev,
                    void *hip_stream);

/* J/K plan over a list of unique quartets (i>=j, k>=l, ij>=kl, each once)
 * whose spherical ERI blocks sit at out_offsets[q] (NULL: packed in list
 * order, the layout of mcint_plan_create) of a device buffer.  Groups the
 * quartets by bra pair and by ket pair once; mcint_jk_plan_execute then
 * overwrites vj_dev / vk_dev (nao x nao row-major, nullable to skip) with
 * J = sum_rs (pq|rs) D_rs and K = sum_qs (pq|rs) D_qs for the symmetric
 * density dm_dev.  No libcint counterpart (PySCF's _vhf direct-SCF
 * contraction plays this role for the reference). */
typedef struct mcint_jk_plan mcint_jk_plan;
int mcint_jk_plan_create(mcint_jk_plan **out, mcint_basis *b, const FINT *shls, size_t nq,
                         const int64_t *out_offsets);
int mcint_jk_plan_execute(mcint_jk_plan *p, const double *eri_dev, const double *dm_dev,
                          double *vj_dev, double *vk_dev, void *hip_stream);
void mcint_jk_plan_destroy(mcint_jk_plan *p);

/* Schwarz bounds q[i*nbas+j] = sqrt(max |(ij|ij)|) over the spherical
 * block of every shell pair among the first nshell shells (nshell <= 0:
 * all).  q is a host array of nbas*nbas doubles. */
int mcint_schwarz(mcint_basis *b, int nshell, double *q);

/* Tuning hook for plan construction (no libcint counterpart; used by the
 * benchmark's autotuner).  MCINT_TUNE_PRIMS_PER_LANE: the team of lanes
 * that shares one contracted quartet is the smallest power of two (<= 64)
 * giving at most this many primitive quartets per lane (default 32).
 * MCINT_TUNE_GENERIC_ONLY: value 1 routes every quartet of plans created
 * afterwards to the generic kernel (runtime class, one workgroup per
 * quartet, no register kernels): the plain path that the benchmark times
 * and cross-checks the register kernels against; 0 (default) restores the
 * class-specialised kernels.  MCINT_TUNE_GRAPH_BRANCHES (1..8, default
 * 4): mcint_plan_execute on a stream captures the plan's buckets into that
 * many concurrent graph branches (buckets balanced by estimated work, one
 * scratch region per branch); takes effect at the next capture.
 * MCINT_TUNE_WAVE_KERNELS (default 1): buckets of the chunked classes run
 * the wave kernel (16-64 lanes per quartet) instead of their
 * component-chunk kernels: 0 never, 1 where measured faster (the tuned
 * table, mcint_tune_class_wave), 2 in every bucket of every class that has
 * one; read at launch.
 * Applies to plans created afterwards; returns the previous value, or a
 * negative status. */
#define MCINT_TUNE_PRIMS_PER_LANE 0
#define MCINT_TUNE_GENERIC_ONLY 1
#define MCINT_TUNE_GRAPH_BRANCHES 2
#define MCINT_TUNE_WAVE_KERNELS 3
int mcint_tune(int key, int value);
/* Per-class override of MCINT_TUNE_PRIMS_PER_LANE for one integral kind
 * (intor code) and canonical class (la >= lb, lc >= ld, (la,lb) >= (lc,ld)
 * for int2e; derivative shell first for ip1; ld = 0 for int3c2e); 0
 * restores the global value.  Written by the autotuner
 * (libcint_amd/autotune.py).  Returns the previous override. */
int mcint_tune_class(int intor, int la, int lb, int lc, int ld, int prims_per_lane);
/* Kernel choice for the buckets of a chunked plain int2e class (canonical
 * order) with teams of `team` lanes (1, 2, 4, .., 64): mode 2 the wave
 * kernel (16-64 lanes per quartet; only where mcint_wave_kernel_exists), 1
 * the component-chunk kernels, 0 the built-in rule (the wave kernel for
 * team-of-one buckets of blocks of 480+ [e0|f0] components).  Applies while
 * MCINT_TUNE_WAVE_KERNELS is 1; read at launch.  Returns the previous mode
 * or a negative status.  (No libcint counterpart: the autotuner's second
 * dimension.) */
int mcint_tune_class_wave(int la, int lb, int lc, int ld, int team, int mode);
int mcint_wave_kernel_exists(int la, int lb, int lc, int ld);

const char *mcint_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MCINT_H */
