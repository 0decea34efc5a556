/*
 * lu2mi_main.c -- single-process driver with lu2's command line and report
 * (lu2.c:1146-1309), running the factorization on the GPU through
 * liblu2mi.so.  The matrix is generated on the device with lu2's drand48
 * sequence (lu2.c:70-85), so it is the same matrix lu2 builds for the seed.
 *
 *   lu2mi [-n N] [-b NB] [-s SEED] [-B board] [-N nboards] [-g] [-p compat|gepp]
 *         [-x xfile] [-h]
 *
 * Timed region as lu2.c:1287-1296: factorization plus back substitution,
 * generation excluded.  Output: the parameter lines (lu2.c:1174-1177),
 * "Emax= " (lu2.c:45-60), "Nswap=... cpsec = ... wsec=... Gflops" with
 * 2n^3/3 flops (lu2.c:1303-1306; Nswap here counts the row exchanges),
 * then the HPL-style residual lines of lu2_mpi.c:2170-2229 and lu2's
 * per-subsystem timer report (timerlib.c:47-70).
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/resource.h>
#include <sys/time.h>
#include "lu2mi.h"

typedef struct {
    int n, seed, nb, boardid, nboards, usehugepage, pivmode;
    const char *xout;
} parms_t;

static void usage(void)
{
    fprintf(stderr, "lu2mi options:\n");
    fprintf(stderr, "  -h: This help\n");
    fprintf(stderr, "  -s: seed (default=1)\n");
    fprintf(stderr, "  -n: size of matrix (default=8192)\n");
    fprintf(stderr, "  -b: block size (default=2048)\n");
    fprintf(stderr, "  -B: board id (default=0): GPU index\n");
    fprintf(stderr, "  -N: number of boards (default=1)\n");
    fprintf(stderr, "  -g: usehugetlbfs (accepted; the matrix lives in GPU memory)\n");
    fprintf(stderr, "  -p: pivot rule compat (lu2's, default) or gepp\n");
    fprintf(stderr, "  -x: file to write the solution x to (n doubles, native binary)\n");
}

static void print_parms(FILE *f, const parms_t *p)
{
    fprintf(f, "N=%d Seed=%d NB=%d usehuge=%d\n", p->n, p->seed, p->nb, p->usehugepage);
    fprintf(f, "Board id=%d # boards=%d\n", p->boardid, p->nboards);
}

static double wall(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

static double cpu(void)
{
    struct rusage ru;
    getrusage(RUSAGE_SELF, &ru);
    return ru.ru_utime.tv_sec + ru.ru_stime.tv_sec + 1e-6 * (ru.ru_utime.tv_usec + ru.ru_stime.tv_usec);
}

/* lu2's per-subsystem report (print_timers, timerlib.c:47-70): the same
 * "name time= ops/cycle=" lines, with time in cycles of the GPU clock and ops
 * the subsystem's algorithmic flops (bytes for the row exchanges), from HIP
 * events around each subsystem's launches.  With the lookahead schedule the
 * panel runs on its own stream beside the trailing update, so the times
 * overlap.  lu2 splits its DGEMM timer by K; here K is always nb. */
static void print_a_timer(const char *name, double ms, double ops, double hz)
{
    const double cyc = ms * 1e-3 * hz;
    printf("%s time=%16.6g ops/cycle=%g\n", name, cyc, cyc > 0 ? ops / cyc : 0.0);
}

static void print_timers(lu2mi_ctx *ctx, int n, int nb)
{
    double t[4 * LU2MI_T_COUNT + 2];
    if (lu2mi_get_subsystem_timing(ctx, t, 4 * LU2MI_T_COUNT + 2) != LU2MI_OK) return;
    const double hz = t[4 * LU2MI_T_COUNT + 1], nd = n;
    const char *dg = nb >= 2048 ? "DGEMM2K    " : nb >= 1024 ? "DGEMM1K    " : nb >= 512 ? "DGEMM512   " : "DGEMMrest  ";
    print_a_timer("swaprows   ", t[4 * LU2MI_T_SWAP], t[4 * LU2MI_T_SWAP + 2], hz);
    print_a_timer("col dec t  ", t[4 * LU2MI_T_PANEL], t[4 * LU2MI_T_PANEL + 1], hz);
    print_a_timer("solve tri  ", t[4 * LU2MI_T_U12], t[4 * LU2MI_T_U12 + 1], hz);
    print_a_timer(dg, t[4 * LU2MI_T_TRAIL], t[4 * LU2MI_T_TRAIL + 1], hz);
    print_a_timer("back sub   ", t[4 * LU2MI_T_BACKSUB], t[4 * LU2MI_T_BACKSUB + 1], hz);
    print_a_timer("Total      ", t[4 * LU2MI_T_COUNT], nd * nd * (nd * 2.0 / 3.0), hz);
}

#define CK(x)                                                                      \
    do {                                                                           \
        int rc_ = (x);                                                             \
        if (rc_) {                                                                 \
            fprintf(stderr, "%s failed (%d): %s\n", #x, rc_, lu2mi_last_error());  \
            exit(-1);                                                              \
        }                                                                          \
    } while (0)

int main(int argc, char *argv[])
{
    parms_t p = {8192, 1, 2048, 0, 1, 0, LU2MI_PIV_COMPAT, NULL};
    static struct option longopts[] = {{"help", no_argument, 0, 'h'},
                                       {"block_size", optional_argument, NULL, 'b'},
                                       {"board_id", optional_argument, NULL, 'B'},
                                       {"nboards", optional_argument, NULL, 'N'},
                                       {"seed", optional_argument, NULL, 's'},
                                       {"ndim_matrix", required_argument, NULL, 'n'},
                                       {"usehugepage", no_argument, 0, 'g'},
                                       {"pivot", required_argument, NULL, 'p'},
                                       {"xout", required_argument, NULL, 'x'},
                                       {NULL, 0, NULL, 0}};
    int ch;
    while ((ch = getopt_long(argc, argv, "B:N:b:ghn:s:p:x:", longopts, NULL)) != -1) {
        switch (ch) {
        case 'b': p.nb = atoi(optarg); break;
        case 'B': p.boardid = atoi(optarg); break;
        case 'N': p.nboards = atoi(optarg); break;
        case 'g': p.usehugepage = 1; break;
        case 's': p.seed = atoi(optarg); break;
        case 'n': p.n = atoi(optarg); break;
        case 'p': p.pivmode = (strcmp(optarg, "gepp") == 0) ? LU2MI_PIV_GEPP : LU2MI_PIV_COMPAT; break;
        case 'x': p.xout = optarg; break;
        case 'h':
        default: usage(); exit(1);
        }
    }
    print_parms(stderr, &p);
    print_parms(stdout, &p);
    const int n = p.n;
    const long lda = (long)n + 16; /* RDIM, lu2.c:28 */
    lu2mi_ctx *ctx = NULL;
    CK(lu2mi_init(p.boardid, &ctx));
    CK(lu2mi_set_option(ctx, LU2MI_OPT_PIVMODE, p.pivmode));
    CK(lu2mi_set_option(ctx, LU2MI_OPT_TIMING, 1));
    double *dA = NULL, *dA0 = NULL, *dx = NULL;
    int *dpiv = NULL;
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)lda;
    CK(lu2mi_malloc(ctx, (void **)&dA, bytes));
    CK(lu2mi_malloc(ctx, (void **)&dx, sizeof(double) * n));
    CK(lu2mi_malloc(ctx, (void **)&dpiv, sizeof(int) * n));
    CK(lu2mi_randomsetmat(ctx, n, p.seed, dA, lda));
    fprintf(stderr, "read/set mat end\n");

    const double w0 = wall(), c0 = cpu();
    CK(lu2mi_lumcolumn_dev(ctx, n, dA, lda, p.nb, dpiv, dx));
    CK(lu2mi_synchronize(ctx));
    const double wtime = wall() - w0, ctime = cpu() - c0;

    /* regenerate A and b (lu2.c:1297-1301) and check */
    double res[10];
    if (lu2mi_malloc(ctx, (void **)&dA0, bytes) == LU2MI_OK) {
        CK(lu2mi_randomsetmat(ctx, n, p.seed, dA0, lda));
        CK(lu2mi_residual_dev(ctx, n, dA0, lda, dx, res));
        lu2mi_free(ctx, dA0);
    } else {
        CK(lu2mi_randomsetmat(ctx, n, p.seed, dA, lda));
        CK(lu2mi_residual_dev(ctx, n, dA, lda, dx, res));
    }
    printf("Emax= %10.3e\n", res[0]);
    int *piv = (int *)malloc(sizeof(int) * n);
    CK(lu2mi_memcpy_d2h(ctx, piv, dpiv, sizeof(int) * n));
    int nswap = 0;
    for (int j = 0; j < n; j++) nswap += (piv[j] != j);
    const double nd = n;
    printf("Nswap=%d cpsec =  %g wsec=%g %g Gflops\n", nswap, ctime, wtime, nd * nd * nd * 2.0 / 3.0 / wtime / 1e9);
    printf("pivot=%s  HPL-2 ||Ax-b||_oo/(eps(||A||_oo||x||_oo+||b||_oo)n) = %.7f (eps=2^-53) %s\n",
           p.pivmode == LU2MI_PIV_GEPP ? "gepp" : "compat", res[6], res[6] < 16 ? "PASSED" : "FAILED");
    printf("||Ax-b||_oo / ( eps * ||A||_1  * N        ) = %16.7f\n", res[7]);
    printf("||Ax-b||_oo / ( eps * ||A||_1  * ||x||_1  ) = %16.7f\n", res[8]);
    printf("||Ax-b||_oo / ( eps * ||A||_oo * ||x||_oo ) = %16.7f\n", res[9]);
    double tm[4];
    CK(lu2mi_get_timing(ctx, tm, 4));
    printf("trailing DGEMM time=%.3f ms launches=%d %.3f TFLOP/s; total %.3f TFLOP/s (2n^3/3+2n^2)\n", tm[0],
           (int)tm[1], tm[0] > 0 ? tm[2] / tm[0] / 1e9 : 0.0, (2.0 * nd * nd * nd / 3.0 + 2.0 * nd * nd) / wtime / 1e12);
    print_timers(ctx, n, p.nb);
    if (p.xout) {
        double *x = (double *)malloc(sizeof(double) * n);
        CK(lu2mi_memcpy_d2h(ctx, x, dx, sizeof(double) * n));
        FILE *f = fopen(p.xout, "wb");
        if (!f || fwrite(x, sizeof(double), (size_t)n, f) != (size_t)n) {
            fprintf(stderr, "cannot write %s\n", p.xout);
            exit(-1);
        }
        fclose(f);
        free(x);
    }
    free(piv);
    lu2mi_free(ctx, dA);
    lu2mi_free(ctx, dx);
    lu2mi_free(ctx, dpiv);
    lu2mi_finalize(ctx);
    return 0;
}
