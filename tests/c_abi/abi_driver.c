/*
 * abi_driver.c -- drives libdimspers_stokes.so through its C ABI alone, as a Julia ccall
 * binding does (MAIN.jl:217-238 -> solveStokes): stk_create -> stk_set_fields -> stk_assemble
 * -> stk_solve -> stk_get_solution -> stk_destroy, then one refused call to check the error
 * path.  Usage: abi_driver DIR METHOD [operators]; DIR holds meta.txt and the raw arrays (native
 * byte order, Julia column-major layouts) written by tests/test_c_abi.py; the solution is written
 * back to DIR/U.bin and DIR/P.bin and the report to DIR/report.txt.
 *
 * With "operators" the driver stands in for StokesPcCG(U, P, KK, MM, GG, Rhs, ifree)
 * (src/IterativeSolvers.jl:94): instead of assembling, it reads the caller's operators as 1-based
 * CSC (K*.bin, M*.bin, G*.bin with the counts in ops.txt, and Rhs.bin), hands them over with
 * stk_set_operators, optionally starts from P0.bin (stk_set_solution: StokesPcCG_pardiso keeps the
 * incoming P) and solves with the options of ops.txt.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dimspers_stokes.h"

static void *load(const char *dir, const char *name, size_t bytes) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    FILE *f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    void *p = malloc(bytes ? bytes : 1);
    if (!p || fread(p, 1, bytes, f) != bytes) { fprintf(stderr, "short read of %s\n", path); exit(2); }
    fclose(f);
    return p;
}

static void save(const char *dir, const char *name, const void *p, size_t bytes) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(p, 1, bytes, f) != bytes) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
    fclose(f);
}

#define CHECK(call)                                                                        \
    do {                                                                                   \
        int rc_ = (call);                                                                  \
        if (rc_ != STK_OK) {                                                               \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, stk_last_error(ctx));     \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

int main(int argc, char **argv) {
    if (argc != 3 && !(argc == 4 && strcmp(argv[3], "operators") == 0)) {
        fprintf(stderr, "usage: %s DIR METHOD [operators]\n", argv[0]);
        return 2;
    }
    const char *dir = argv[1];
    const int method = atoi(argv[2]);
    const int operators = argc == 4;
    char path[4096];
    snprintf(path, sizeof path, "%s/meta.txt", dir);
    FILE *m = fopen(path, "r");
    long long nel, nnod, nvert, nfixed;
    double g;
    if (!m || fscanf(m, "%lld %lld %lld %lld %lf", &nel, &nnod, &nvert, &nfixed, &g) != 5) {
        fprintf(stderr, "bad meta.txt\n");
        return 2;
    }
    fclose(m);
    const int64_t *e2n = load(dir, "e2n.bin", 6 * (size_t)nel * 8);
    const double *th = load(dir, "el_theta.bin", 3 * (size_t)nel * 8);
    const double *r = load(dir, "el_r.bin", 3 * (size_t)nel * 8);
    const double *nth = load(dir, "node_theta.bin", (size_t)nnod * 8);
    const double *nr = load(dir, "node_r.bin", (size_t)nnod * 8);
    const int64_t *fixed = load(dir, "fixed.bin", (size_t)nfixed * 8);
    const double *rho = load(dir, "rho.bin", (size_t)nnod * 8);
    const double *eta = load(dir, "eta.bin", 6 * (size_t)nel * 8);
    const double *C[6];
    for (int k = 0; k < 6; ++k) {
        char name[32];
        snprintf(name, sizeof name, "C%d.bin", k);
        C[k] = load(dir, name, 6 * (size_t)nel * 8);
    }

    stk_ctx *ctx = NULL;
    if (stk_create(&ctx, nel, nnod, nvert, e2n, th, r, nth, nr, fixed, nfixed) != STK_OK) {
        fprintf(stderr, "stk_create failed: %s\n", stk_last_error(NULL));
        return 1;
    }
    int64_t info[7];
    CHECK(stk_info(ctx, info));
    stk_opts o;
    stk_default_opts(&o);
    o.method = method;
    o.stop = STK_STOP_FULL;
    o.rtol = 1e-8;
    if (!operators) {
        CHECK(stk_set_fields(ctx, g, rho, eta, 6, C));
        CHECK(stk_assemble(ctx, STK_ASM_FP64 | STK_ASM_FP32));
    } else {
        long long nK, nM, nG;
        int k_lower, m_lower, warm, ppc, stop, null_mode, itmax;
        double rtol, inner_rtol;
        snprintf(path, sizeof path, "%s/ops.txt", dir);
        FILE *f = fopen(path, "r");
        if (!f || fscanf(f, "%lld %lld %lld %d %d %d %d %d %d %d %lf %lf", &nK, &nM, &nG, &k_lower, &m_lower, &warm,
                         &ppc, &stop, &null_mode, &itmax, &rtol, &inner_rtol) != 12) {
            fprintf(stderr, "bad ops.txt\n");
            return 2;
        }
        fclose(f);
        const size_t nU = (size_t)info[3], nP = (size_t)info[4];
        const int64_t *Kc = load(dir, "Kcolptr.bin", (nU + 1) * 8), *Kr = load(dir, "Krowval.bin", (size_t)nK * 8);
        const double *Kv = load(dir, "Knzval.bin", (size_t)nK * 8);
        const int64_t *Mc = load(dir, "Mcolptr.bin", (nP + 1) * 8), *Mr = load(dir, "Mrowval.bin", (size_t)nM * 8);
        const double *Mv = load(dir, "Mnzval.bin", (size_t)nM * 8);
        const int64_t *Gc = load(dir, "Gcolptr.bin", (nP + 1) * 8), *Gr = load(dir, "Growval.bin", (size_t)nG * 8);
        const double *Gv = load(dir, "Gnzval.bin", (size_t)nG * 8);
        const double *Rhs = load(dir, "Rhs.bin", nU * 8);
        CHECK(stk_set_operators(ctx, k_lower, Kc, Kr, Kv, m_lower, Mc, Mr, Mv, Gc, Gr, Gv, Rhs, STK_ASM_FP64));
        if (warm) {
            const double *P0 = load(dir, "P0.bin", nP * 8);
            CHECK(stk_set_solution(ctx, NULL, P0));
        }
        o.inner = STK_INNER_JACOBI;
        o.inner_itmax = 100000;
        o.warm_start = warm;
        o.pressure_pc = ppc;
        o.stop = stop;
        o.null_mode = null_mode;
        o.itmax = itmax;
        o.rtol = rtol;
        o.inner_rtol = inner_rtol;
    }
    stk_report rep;
    CHECK(stk_solve(ctx, &o, &rep));
    double *U = malloc((size_t)info[3] * 8), *P = malloc((size_t)info[4] * 8);
    CHECK(stk_get_solution(ctx, U, P));
    save(dir, "U.bin", U, (size_t)info[3] * 8);
    save(dir, "P.bin", P, (size_t)info[4] * 8);

    /* an unknown method is refused with STK_E_ARG and a message, and the context stays usable */
    stk_opts bad = o;
    bad.method = 9;
    const int rc_bad = stk_solve(ctx, &bad, NULL);
    char msg[512];
    const char *last = stk_last_error(ctx);   /* valid until the next call on ctx: copy it */
    snprintf(msg, sizeof msg, "%s", last ? last : "");
    CHECK(stk_get_solution(ctx, U, P));

    snprintf(path, sizeof path, "%s/report.txt", dir);
    FILE *out = fopen(path, "w");
    if (!out) return 2;
    fprintf(out, "%d %d %lld %.17g %d %s\n", rep.converged, rep.outer_iterations, (long long)rep.inner_iterations,
            rep.full_residual, rc_bad, msg[0] ? "message" : "empty");
    fclose(out);
    stk_destroy(ctx);
    printf("[abi_driver] nU %lld nP %lld converged %d outer %d residual %.3e\n", (long long)info[3],
           (long long)info[4], rep.converged, rep.outer_iterations, rep.full_residual);
    return 0;
}
