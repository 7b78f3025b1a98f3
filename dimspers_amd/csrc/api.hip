// extern "C" entry points of libdimspers_stokes.so (declared in include/dimspers_stokes.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>


#include "device_ops.h"
#include "fabric.h"
#include "mlp.h"
#include "solver.h"
#include "stress.h"
#include "thermal.h"

namespace stk {
template <typename S, typename D> void convert(Context &c, const S *src, D *dst, int64_t n);
template <typename T> void lumped_inverse(Context &c, const Pattern &p, const T *M, T *pc);
template <typename T> void csr_diag_inverse(Context &c, const Pattern &p, const T *A, T *dinv);
void host_spmv(Context &c, int which, int prec, const double *xh, double *yh);
double update_velocity(Context &c, double *Ucart_dev);

Context::~Context() {
    delete mlp;
    if (h_pub) (void)hipHostFree(h_pub);
    if (h_seq) (void)hipHostFree(h_seq);
    if (dense_ev) (void)hipEventDestroy(dense_ev);
    if (dense_ev2) (void)hipEventDestroy(dense_ev2);
    if (dense_stream) (void)hipStreamDestroy(dense_stream);
    if (aux_stream) (void)hipStreamDestroy(aux_stream);
    if (stream) (void)hipStreamDestroy(stream);
}
}  // namespace stk

struct stk_ctx {
    stk::Context c;
};

using namespace stk;

static thread_local std::string g_err;

template <typename F>
static int guarded(stk_ctx *ctx, F f) {
    try {
        if (ctx) {
            STK_HIP(hipSetDevice(ctx->c.device));
            ctx->c.err.clear();
        }
        return f();
    } catch (const Error &e) {
        (ctx ? ctx->c.err : g_err) = e.what();
        return e.code;
    } catch (const std::exception &e) {
        (ctx ? ctx->c.err : g_err) = e.what();
        return STK_E_HIP;
    }
}

extern "C" {

void stk_default_opts(stk_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->method = STK_SOLVE_PLAIN;
    o->inner = STK_INNER_MG;
    o->stop = STK_STOP_SCHUR;
    o->itmax = 200;
    o->itmin = 10;
    o->inner_itmax = 2000;
    o->refine_max = 30;
    o->cheb_degree = 2;
    o->rtol = 1e-8;
    o->inner_rtol = 1e-11;
    o->mixed_inner_rtol = 0.2;
    o->mixed_schur_rtol = 0.1;
    o->stall_ratio = 0.5;
    o->null_mode = STK_NULL_PROJECT;
    o->pc_sweeps = 3;
    o->fine_smoother = STK_SMOOTH_GS;
}

const char *stk_last_error(const stk_ctx *ctx) { return ctx ? ctx->c.err.c_str() : g_err.c_str(); }

void stk_destroy(stk_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    delete ctx;
}

int stk_create(stk_ctx **out, int64_t nel, int64_t nnod, int64_t nvert, const int64_t *e2n,
               const double *el_theta, const double *el_r, const double *node_theta, const double *node_r,
               const int64_t *fixed_dofs, int64_t nfixed) {
    g_err.clear();
    if (!out) { g_err = "ctx pointer is NULL"; return STK_E_ARG; }
    *out = nullptr;
    stk_ctx *ctx = new stk_ctx();
    int rc = guarded(nullptr, [&]() {
        Context &c = ctx->c;
        STK_CHECK(nel > 0 && nnod > 0 && nvert > 0 && nvert <= nnod, STK_E_SHAPE, "bad mesh sizes");
        STK_CHECK(nnod < (int64_t(1) << 30), STK_E_SHAPE, "mesh too large for 32-bit node ids");
        STK_CHECK(e2n && el_theta && el_r && node_theta && node_r, STK_E_ARG, "NULL mesh array");
        STK_CHECK(nfixed >= 0, STK_E_ARG, "negative number of fixed dofs");
        STK_CHECK(nfixed == 0 || fixed_dofs, STK_E_ARG, "fixed_dofs is NULL");
        STK_HIP(hipGetDevice(&c.device));
        STK_HIP(hipStreamCreate(&c.stream));
        STK_HIP(hipStreamCreate(&c.aux_stream));
        STK_HIP(hipStreamCreate(&c.dense_stream));
        STK_HIP(hipEventCreateWithFlags(&c.dense_ev, hipEventDisableTiming));
        STK_HIP(hipEventCreateWithFlags(&c.dense_ev2, hipEventDisableTiming));
        c.nel = nel; c.nnod = nnod; c.nvert = nvert; c.nU = 2 * nnod; c.nP = nvert;
        c.e2n.resize(6 * nel);
        for (int64_t e = 0; e < nel; ++e)
            for (int a = 0; a < 6; ++a) {
                const int64_t v = e2n[e * 6 + a] - 1;
                STK_CHECK(v >= 0 && v < nnod, STK_E_SHAPE, "e2n entry out of range");
                if (a < 3) STK_CHECK(v < nvert, STK_E_SHAPE, "vertex id beyond nvert");
                c.e2n[e * 6 + a] = (int32_t)v;
            }
        c.el_th.assign(el_theta, el_theta + 3 * nel);
        c.el_r.assign(el_r, el_r + 3 * nel);
        c.node_th.assign(node_theta, node_theta + nnod);
        c.node_r.assign(node_r, node_r + nnod);
        c.fixed.assign(2 * nnod, 0);
        for (int64_t i = 0; i < nfixed; ++i) {
            const int64_t d = fixed_dofs[i] - 1;
            STK_CHECK(d >= 0 && d < 2 * nnod, STK_E_SHAPE, "fixed dof out of range");
            c.fixed[d] = 1;
        }
        for (int64_t d = 0; d < 2 * nnod; ++d)
            if (c.fixed[d]) c.fixed_list.push_back(d);
        c.d_fixlist.upload(c.fixed_list, c.stream);
        build_patterns(c);
        build_colouring(c);
        build_slot_maps(c);
        detect_structured(c);
        // The rigid rotation (u_theta = r, u_r = 0) satisfies the boundary conditions, and is then
        // a near-null mode of K, exactly when no tangential dof is fixed.
        c.rotation_free = true;
        for (int64_t d : c.fixed_list)
            if (d % 2 == 0) c.rotation_free = false;
        // The multigrid's coarse constraints are those of free slip on both walls (the radial dof
        // of every node on the smallest and largest radius, free_slip in
        // src/Setup/boundary_conditions.jl:27-37); other fixed sets solve without it.
        if (!c.structured)
            c.mg_note = "the multigrid needs a structured annulus mesh (Grid_split1 or Grid)";
        else if (!fixed_set_is_free_slip(c))
            c.mg_note = "the multigrid needs free slip on both walls (the fixed set differs)";
        else
            build_mg_structure(c);
        if (!c.mg.empty()) {
            build_gs_layout(c);
            c.gs.on = true;
        }
        stress_prepare(c);           // natural-order element data (assembly, stress, statistics)
        if (!c.mg.empty()) {
            STK_CHECK(c.mg[0].pat.ptr == c.MN.ptr && c.mg[0].pat.col == c.MN.col, STK_E_SHAPE,
                      "vertex pattern mismatch between mesh and multigrid level 0");
            for (MGLevel &L : c.mg) {
                std::vector<int64_t> fl;
                for (int64_t d = 0; d < 2 * L.nvert; ++d)
                    if (L.h_fixed[d]) fl.push_back(d);
                L.nfix = (int64_t)fl.size();
                L.fixlist.upload(fl, c.stream);
            }
        }
        // reduction partials: the largest node-row grid of any SpMV with a fused dot
        const int64_t maxrows = std::max<int64_t>(c.nnod, c.nvert);
        c.red64.alloc(std::max<int64_t>(2 * kReduceBlocks, 2 * (maxrows / 32 + 64)));   // up to 2 dots
        c.d_sc.alloc(SC_COUNT);
        c.d_sc.zero(c.stream);
        STK_HIP(hipHostMalloc(&c.h_pub, SC_COUNT * sizeof(double), hipHostMallocCoherent));
        STK_HIP(hipHostMalloc(&c.h_seq, 256, hipHostMallocCoherent));
        *(volatile unsigned *)c.h_seq = 0u;
        // statistics data (stk_stats, stk_stats_noalloc) once per mesh; a mesh without the radius
        // levels of write_stats keeps the reason for those calls
        try {
            statistics_prepare(c);
        } catch (const Error &e) {
            c.stats_note = e.what();
        }
        // this device's copy of stk_stats_noalloc's scratch array (the device is current)
        c.stats_scratch = stats_scratch_address(c.device);
        STK_HIP(hipStreamSynchronize(c.stream));
        return STK_OK;
    });
    if (rc != STK_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return STK_OK;
}

int stk_info(const stk_ctx *ctx, int64_t *o) {
    if (!ctx || !o) return STK_E_ARG;
    const Context &c = ctx->c;
    o[0] = c.nel; o[1] = c.nnod; o[2] = c.nvert; o[3] = c.nU; o[4] = c.nP;
    o[5] = (int64_t)c.col_off.size() - 1;
    o[6] = (int64_t)c.mg.size();
    return STK_OK;
}

int stk_colouring(stk_ctx *ctx, int32_t which, int32_t *colour, int32_t *ncolour) {
    if (!ctx || !colour || !ncolour) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        if (which == STK_COLOUR_ELEMENTS) {
            const int nc = (int)c.col_off.size() - 1;
            for (int k = 0; k < nc; ++k)
                for (int64_t p = c.col_off[k]; p < c.col_off[k + 1]; ++p) colour[c.perm[p]] = k;
            *ncolour = nc;
        } else if (which == STK_COLOUR_NODES) {
            STK_CHECK(c.gs.on, STK_E_STATE, "no Gauss-Seidel colouring (mesh without the multigrid)");
            for (int64_t i = 0; i < c.nnod; ++i) colour[i] = c.gs.node_colour[i];
            *ncolour = c.gs.ncolour;
        } else {
            throw Error(STK_E_ARG, "unknown colouring");
        }
        return STK_OK;
    });
}

int stk_set_fields(stk_ctx *ctx, double g, const double *rho, const double *eta_ip, int64_t nip,
                   const double *const C[6]) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(rho && eta_ip, STK_E_ARG, "NULL field array");
        STK_CHECK(nip >= 6, STK_E_SHAPE, "eta.ip needs at least 6 integration-point columns");
        c.g = g;
        c.d_rho.upload(rho, c.nnod, c.stream);
        // Julia nel x nip column-major: the first columns are contiguous; up to 7 are kept (the
        // 6-point rule reads 6, the 7-point rule 7)
        const int64_t ncol = std::min<int64_t>(nip, 7);
        c.field_cols = ncol;
        c.d_eta.upload(eta_ip, ncol * c.nel, c.stream);
        if (C) {
            std::vector<double> buf(6 * ncol * c.nel);
            for (int k = 0; k < 6; ++k) {
                STK_CHECK(C[k], STK_E_ARG, "NULL tensor component");
                std::copy(C[k], C[k] + ncol * c.nel, buf.begin() + (int64_t)k * ncol * c.nel);
            }
            c.d_C.upload(buf, c.stream);
            c.have_C = true;
            STK_HIP(hipStreamSynchronize(c.stream));
        } else {
            c.have_C = false;
        }
        STK_HIP(hipStreamSynchronize(c.stream));
        c.have_fields = true;
        return STK_OK;
    });
}

}  // extern "C"

namespace {

template <typename T>
void assemble_prec(Context &c, int flags, bool mg, DBuf<T> &K, DBuf<T> &G, DBuf<T> &GT, DBuf<T> &M, DBuf<T> &F,
                   DBuf<T> &RK, bool f64) {
    // the gathers of the element assembly write every stored position of these
    K.alloc(4 * c.KN.nstore());
    G.alloc(2 * c.GN.nstore());
    GT.alloc(2 * c.GTN.nstore());
    M.alloc(c.MN.nnz());
    F.alloc(c.nU);
    T *A1 = nullptr, *E1 = nullptr;
    auto Abuf = [&](MGLevel &L) -> DBuf<T> & {
        if constexpr (sizeof(T) == 8) return L.A64; else return L.A32;
    };
    auto Ebuf = [&](MGLevel &L) -> DBuf<T> & {
        if constexpr (sizeof(T) == 8) return L.E64; else return L.E32;
    };
    if (mg) {
        for (size_t l = 0; l < c.mg.size(); ++l) {
            MGLevel &L = c.mg[l];
            Abuf(L).alloc(4 * L.pat.nstore());
            if (l > 0) Abuf(L).zero(c.stream);   // the coarse Galerkin levels accumulate; level 0 is gathered
            Ebuf(L).alloc(36 * L.nel);
        }
        A1 = Abuf(c.mg[0]).p;
        E1 = Ebuf(c.mg[0]).p;
        RK.alloc(4 * c.GTN.nstore());
    } else {
        RK.release();
    }
    assemble_impl<T>(c, flags, K.p, G.p, GT.p, M.p, F.p, A1, E1, mg ? RK.p : nullptr);
    if (mg)
        for (size_t l = 0; l + 1 < c.mg.size(); ++l)
            galerkin_impl<T>(c, l, Ebuf(c.mg[l]).p, Ebuf(c.mg[l + 1]).p, Abuf(c.mg[l + 1]).p);
    if (!(flags & STK_ASM_NO_BC)) fix_diag<T>(c, c.KN, (int64_t)c.fixed_list.size(), c.d_fixlist.p, K.p);
    if (mg)
        for (MGLevel &L : c.mg) fix_diag<T>(c, L.pat, L.nfix, L.fixlist.p, Abuf(L).p);
    (void)f64;
}

// The lumped preconditioner of the FP64 operators and, with fp32, their rounded copies.
void derive_from_fp64(Context &c, bool fp32) {
    c.pc64.alloc(c.nP);
    lumped_inverse<double>(c, c.MN, c.M64.p, c.pc64.p);
    if (fp32) {
        c.K32.alloc(c.K64.n); convert<double, float>(c, c.K64.p, c.K32.p, c.K64.n);
        c.G32.alloc(c.G64.n); convert<double, float>(c, c.G64.p, c.G32.p, c.G64.n);
        c.GT32.alloc(c.GT64.n); convert<double, float>(c, c.GT64.p, c.GT32.p, c.GT64.n);
        c.M32.alloc(c.M64.n); convert<double, float>(c, c.M64.p, c.M32.p, c.M64.n);
        c.F32.alloc(c.F64.n); convert<double, float>(c, c.F64.p, c.F32.p, c.F64.n);
        c.pc32.alloc(c.nP); convert<double, float>(c, c.pc64.p, c.pc32.p, c.nP);
    }
}

}  // namespace

extern "C" {

int stk_assemble(stk_ctx *ctx, int32_t flags) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.have_fields, STK_E_STATE, "stk_set_fields must precede stk_assemble");
        STK_CHECK(!(flags & STK_ASM_7POINT) || c.field_cols >= 7, STK_E_STATE,
                  "the 7-point rule needs 7 columns of eta.ip and of the tensor in stk_set_fields");
        if (!c.gK.npos) build_gather_lists(c);   // once per mesh, at its first assembly
        const bool fp32 = flags & STK_ASM_FP32;
        const bool fp64 = (flags & STK_ASM_FP64) || !fp32;
        const bool solve_form = !(flags & (STK_ASM_CARTESIAN | STK_ASM_NO_BC));
        const bool mg = solve_form && c.structured && !(flags & STK_ASM_NO_MG) && !c.mg.empty();
        c.assembled = false;
        c.loaded = false;
        c.mg_built = false;
        c.coarse64_valid = false;
        c.have_mg64 = c.have_mg32 = false;
        // stale products of an earlier assembly
        c.K32.release(); c.G32.release(); c.GT32.release(); c.M32.release(); c.F32.release(); c.pc32.release();
        c.RK32.release();
        for (MGLevel &L : c.mg) { L.RA64.release(); L.RA32.release(); }   // rebuilt by the solver set-up
        c.S64.release(); c.S32.release(); c.S32s.release(); c.Sdinv64.release(); c.Sdinv32.release();
        c.pic.valid = false;
        if (fp64) {
            assemble_prec<double>(c, flags, mg, c.K64, c.G64, c.GT64, c.M64, c.F64, c.RK64, true);
            derive_from_fp64(c, fp32);
            if (fp32) {
                if (mg) {
                    for (MGLevel &L : c.mg) {
                        L.A32.alloc(L.A64.n);
                        convert<double, float>(c, L.A64.p, L.A32.p, L.A64.n);
                    }
                    c.RK32.alloc(c.RK64.n);
                    convert<double, float>(c, c.RK64.p, c.RK32.p, c.RK64.n);
                }
            }
        } else {
            c.K64.release(); c.G64.release(); c.GT64.release(); c.M64.release(); c.pc64.release(); c.RK64.release();
            assemble_prec<float>(c, flags, mg, c.K32, c.G32, c.GT32, c.M32, c.F32, c.RK32, false);
            c.pc32.alloc(c.nP);
            lumped_inverse<float>(c, c.MN, c.M32.p, c.pc32.p);
            c.F64.alloc(c.nU);
            convert<float, double>(c, c.F32.p, c.F64.p, c.nU);
        }
        STK_HIP(hipStreamSynchronize(c.stream));
        c.mg_built = mg;
        c.asm_flags = flags;
        if (solve_form) {
            if (fp64 && fp32 && mg) {
                setup_fast(c);
            } else {
                if (fp64) setup_solver_data<double>(c);
                if (fp32) setup_solver_data<float>(c);
            }
        }
        c.asm_flags = flags;
        c.assembled = true;
        STK_HIP(hipStreamSynchronize(c.stream));
        return STK_OK;
    });
}

int stk_export_csc(stk_ctx *ctx, int32_t which, int32_t lower_only, int64_t *nrows, int64_t *ncols, int64_t *nnz,
                   int64_t *colptr, int64_t *rowval, double *nzval) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        if (which == STK_MAT_COARSE_INV) {
            // dense and symmetric: column j is row j of the row-major array
            STK_CHECK(c.mg_built && c.coarse64_valid && c.coarse_n > 0, STK_E_STATE, "no coarsest inverse");
            const int64_t n = c.coarse_n;
            if (nrows) *nrows = n;
            if (ncols) *ncols = n;
            if (nnz) *nnz = n * n;
            if (!colptr || !rowval || !nzval) return (int)STK_OK;
            STK_HIP(hipMemcpyAsync(nzval, c.coarse_inv64.p, (size_t)(n * n) * sizeof(double), hipMemcpyDeviceToHost,
                                   c.stream));
            STK_HIP(hipStreamSynchronize(c.stream));
            for (int64_t j = 0; j <= n; ++j) colptr[j] = j * n + 1;
            for (int64_t j = 0; j < n; ++j)
                for (int64_t i = 0; i < n; ++i) rowval[j * n + i] = i + 1;
            return (int)STK_OK;
        }
        const bool thermal = which == STK_MAT_KT || which == STK_MAT_MT;
        STK_CHECK(thermal ? c.have_thermal : c.assembled, STK_E_STATE, "nothing assembled");
        const Pattern *p = nullptr;
        int br = 1, bc = 1;
        int64_t nr = 0, nc = 0;
        const double *v64 = nullptr;
        const float *v32 = nullptr;
        switch (which) {
            case STK_MAT_K: p = &c.KN; br = bc = 2; nr = nc = c.nU; v64 = c.K64.p; v32 = c.K32.p; break;
            case STK_MAT_G: p = &c.GN; br = 2; nr = c.nU; nc = c.nP; v64 = c.G64.p; v32 = c.G32.p; break;
            case STK_MAT_GT: p = &c.GTN; bc = 2; nr = c.nP; nc = c.nU; v64 = c.GT64.p; v32 = c.GT32.p; break;
            case STK_MAT_M: p = &c.MN; nr = nc = c.nP; v64 = c.M64.p; v32 = c.M32.p; break;
            case STK_MAT_S: p = &c.MN; nr = nc = c.nP; v64 = c.S64.p; v32 = c.S32.p; break;
            case STK_MAT_KT: p = &c.KN; nr = nc = c.nnod; v64 = c.KT.p; break;
            case STK_MAT_MT: p = &c.KN; nr = nc = c.nnod; v64 = c.MT.p; break;
            case STK_MAT_R:
                STK_CHECK(c.mg_built, STK_E_STATE, "no multigrid hierarchy");
                p = &c.mg[0].R_pat; br = bc = 2; nr = 2 * c.mg[0].nvert; nc = c.nU;
                v64 = c.mg[0].R64.p; v32 = c.mg[0].R32.p;
                break;
            case STK_MAT_RK:
                STK_CHECK(c.mg_built, STK_E_STATE, "no multigrid hierarchy");
                p = &c.GTN; br = bc = 2; nr = 2 * c.nP; nc = c.nU; v64 = c.RK64.p; v32 = c.RK32.p;
                break;
            default: {
                // multigrid levels and prolongations (set-up checks of the hierarchy)
                const bool level = which >= STK_MAT_LEVEL && which < STK_MAT_PROLONG;
                const bool prol = which >= STK_MAT_PROLONG && which < STK_MAT_PROLONG + 32;
                const bool ra = which >= STK_MAT_LEVEL_RA && which < STK_MAT_LEVEL_RA + 32;
                if (!level && !prol && !ra) throw Error(STK_E_ARG, "unknown operator");
                STK_CHECK(c.mg_built, STK_E_STATE, "no multigrid hierarchy");
                const int64_t l = which - (level ? STK_MAT_LEVEL : prol ? STK_MAT_PROLONG : STK_MAT_LEVEL_RA);
                STK_CHECK(l < (int64_t)c.mg.size() && (!ra || l >= 1), STK_E_ARG, "no such multigrid level");
                MGLevel &L = c.mg[l];
                br = bc = 2;
                if (level) {
                    p = &L.pat; nr = nc = 2 * L.nvert; v64 = L.A64.p; v32 = L.A32.p;
                } else if (ra) {
                    p = &L.RA_pat; nr = 2 * L.nvert; nc = 2 * c.mg[l - 1].nvert; v64 = L.RA64.p; v32 = L.RA32.p;
                } else {
                    p = &L.P_pat; nr = l == 0 ? c.nU : 2 * c.mg[l - 1].nvert; nc = 2 * L.nvert;
                    v64 = L.P64.p; v32 = L.P32.p;
                }
            }
        }
        STK_CHECK(v64 || v32, STK_E_STATE, "operator not available");
        const int64_t nvals = p->nstore() * br * bc;
        std::vector<double> vals(nvals);
        if (v64) {
            STK_HIP(hipMemcpyAsync(vals.data(), v64, nvals * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        } else {
            std::vector<float> f(nvals);
            STK_HIP(hipMemcpyAsync(f.data(), v32, nvals * sizeof(float), hipMemcpyDeviceToHost, c.stream));
            STK_HIP(hipStreamSynchronize(c.stream));
            std::copy(f.begin(), f.end(), vals.begin());
        }
        STK_HIP(hipStreamSynchronize(c.stream));
        // count entries per column
        std::vector<int64_t> cnt(nc + 1, 0);
        auto visit = [&](auto fn) {
            for (int64_t i = 0; i < p->nrows; ++i)
                for (int32_t k = p->ptr[i]; k < p->ptr[i + 1]; ++k) {
                    const int64_t q0 = p->pos(i, k - p->ptr[i]);
                    for (int r = 0; r < br; ++r)
                        for (int q = 0; q < bc; ++q) {
                            const int64_t row = i * br + r, col = (int64_t)p->col[k] * bc + q;
                            if (lower_only && row < col) continue;
                            fn(row, col, vals[q0 * br * bc + r * bc + q]);
                        }
                }
        };
        visit([&](int64_t, int64_t col, double) { cnt[col + 1]++; });
        for (int64_t j = 0; j < nc; ++j) cnt[j + 1] += cnt[j];
        if (nrows) *nrows = nr;
        if (ncols) *ncols = nc;
        if (nnz) *nnz = cnt[nc];
        if (!colptr || !rowval || !nzval) return (int)STK_OK;
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        // rows are visited in increasing order, so each column comes out sorted
        visit([&](int64_t row, int64_t col, double v) {
            const int64_t d = pos[col]++;
            rowval[d] = row + 1;
            nzval[d] = v;
        });
        for (int64_t j = 0; j <= nc; ++j) colptr[j] = cnt[j] + 1;
        return (int)STK_OK;
    });
}

int stk_export_rhs(stk_ctx *ctx, double *F) {
    if (!ctx || !F) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.assembled && c.F64, STK_E_STATE, "nothing assembled");
        c.F64.download(F, c.nU, c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));
        return STK_OK;
    });
}

int stk_spmv(stk_ctx *ctx, int32_t which, int32_t precision, const double *x, double *y) {
    if (!ctx || !x || !y) return STK_E_ARG;
    return guarded(ctx, [&]() {
        STK_CHECK(ctx->c.assembled, STK_E_STATE, "nothing assembled");
        host_spmv(ctx->c, which, precision, x, y);
        return STK_OK;
    });
}

int stk_solve(stk_ctx *ctx, const stk_opts *opts, stk_report *report) {
    if (!ctx) return STK_E_ARG;
    stk_report local{};
    stk_report &rep = report ? *report : local;
    std::memset(&rep, 0, sizeof(rep));
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.assembled, STK_E_STATE, "stk_assemble must precede stk_solve");
        STK_CHECK(!(c.asm_flags & (STK_ASM_CARTESIAN | STK_ASM_NO_BC)), STK_E_STATE,
                  "the solve needs the rotated operators with boundary conditions");
        stk_opts o;
        stk_default_opts(&o);
        if (opts) o = *opts;
        STK_CHECK(o.method >= STK_SOLVE_PLAIN && o.method <= STK_SOLVE_LEARNED, STK_E_ARG, "bad method");
        STK_CHECK(o.inner == STK_INNER_MG || o.inner == STK_INNER_JACOBI, STK_E_ARG, "bad inner solver");
        STK_CHECK(o.itmax > 0 && o.inner_itmax > 0 && o.rtol > 0 && o.cheb_degree >= 1, STK_E_ARG, "bad options");
        STK_CHECK(o.fine_smoother == STK_SMOOTH_GS || o.fine_smoother == STK_SMOOTH_CHEB, STK_E_ARG, "bad smoother");
        STK_CHECK(o.warm_start == 0 || o.warm_start == 1, STK_E_ARG, "bad warm_start");
        STK_CHECK(o.pressure_pc == STK_PPC_LUMPED || o.pressure_pc == STK_PPC_IC, STK_E_ARG, "bad pressure_pc");
        STK_CHECK(o.pressure_pc == STK_PPC_LUMPED || o.method == STK_SOLVE_PLAIN, STK_E_ARG,
                  "the incomplete Cholesky pressure preconditioner belongs to the plain FP64 method");
        // the FP32 methods stop their inner solves at a loose tolerance that only the multigrid
        // preconditioner turns into progress; with point-Jacobi inner solves their sweeps stall
        STK_CHECK(o.method == STK_SOLVE_PLAIN || o.inner == STK_INNER_MG, STK_E_ARG,
                  "the mixed and learned methods need the multigrid inner solver");
        c.fine_smoother = o.fine_smoother;
        if (o.method == STK_SOLVE_PLAIN) STK_CHECK(c.K64, STK_E_STATE, "plain solve needs FP64 operators");
        const int rc = run_solve(c, o, rep);
        if (rc == STK_E_NOTCONV) c.err = "stopped at the iteration cap before the tolerance";
        return rc;
    });
}

int stk_set_solution(stk_ctx *ctx, const double *U, const double *P) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        c.U.alloc(c.nU);
        c.P.alloc(c.nP);
        if (U) c.U.upload(U, c.nU, c.stream); else c.U.zero(c.stream);
        if (P) c.P.upload(P, c.nP, c.stream); else c.P.zero(c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));
        return STK_OK;
    });
}

int stk_set_operators(stk_ctx *ctx, int32_t K_lower, const int64_t *K_colptr, const int64_t *K_rowval,
                      const double *K_nzval, int32_t M_lower, const int64_t *M_colptr, const int64_t *M_rowval,
                      const double *M_nzval, const int64_t *G_colptr, const int64_t *G_rowval,
                      const double *G_nzval, const double *Rhs, int32_t flags) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK((flags & ~(STK_ASM_FP64 | STK_ASM_FP32)) == 0, STK_E_ARG,
                  "stk_set_operators takes STK_ASM_FP64 and STK_ASM_FP32 only");
        const bool fp32 = flags & STK_ASM_FP32;
        std::vector<double> K, G, GT, M, F;
        import_operators(c, K_lower != 0, CscIn{K_colptr, K_rowval, K_nzval}, M_lower != 0,
                         CscIn{M_colptr, M_rowval, M_nzval}, CscIn{G_colptr, G_rowval, G_nzval}, Rhs, K, G, GT, M,
                         F);
        c.assembled = false;
        c.mg_built = false;
        c.coarse64_valid = false;
        c.have_mg64 = c.have_mg32 = false;
        c.pic.valid = false;
        c.K32.release(); c.G32.release(); c.GT32.release(); c.M32.release(); c.F32.release(); c.pc32.release();
        c.RK64.release(); c.RK32.release();
        c.S64.release(); c.S32.release(); c.S32s.release(); c.Sdinv64.release(); c.Sdinv32.release();
        c.K64.upload(K, c.stream);
        c.G64.upload(G, c.stream);
        c.GT64.upload(GT, c.stream);
        c.M64.upload(M, c.stream);
        c.F64.upload(F, c.stream);
        fix_diag<double>(c, c.KN, (int64_t)c.fixed_list.size(), c.d_fixlist.p, c.K64.p);
        derive_from_fp64(c, fp32);
        STK_HIP(hipStreamSynchronize(c.stream));
        // the caller's operators, not this library's assembly: no multigrid (it is built from the
        // element matrices of stk_assemble)
        c.loaded = true;
        c.asm_flags = STK_ASM_FP64 | (fp32 ? STK_ASM_FP32 : 0);
        setup_solver_data<double>(c);
        if (fp32) setup_solver_data<float>(c);
        c.assembled = true;
        STK_HIP(hipStreamSynchronize(c.stream));
        return STK_OK;
    });
}

int stk_get_solution(stk_ctx *ctx, double *U, double *P) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.U && c.P, STK_E_STATE, "no solution");
        if (U) c.U.download(U, c.nU, c.stream);
        if (P) c.P.download(P, c.nP, c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));
        return STK_OK;
    });
}

int stk_time_kernel(stk_ctx *ctx, int32_t kernel, int32_t reps, double *avg_ms, double *bytes) {
    if (!ctx || !avg_ms || !bytes || reps <= 0) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        const bool standalone = kernel == STK_KERNEL_FSE || kernel == STK_KERNEL_ANISO || kernel == STK_KERNEL_COPY;
        STK_CHECK(standalone || c.assembled, STK_E_STATE, "nothing assembled");
        if (kernel == STK_KERNEL_ASSEMBLY) {
            STK_CHECK(c.have_fields && c.K64, STK_E_STATE, "FP64 assembly needs fields");
            const int flags = c.asm_flags;
            EventPair ev_;
            hipEvent_t e0 = ev_.a, e1 = ev_.b;
            // element kernels only (all colours), into the FP64 operators of the last assembly
            const bool mg = c.mg_built;
            STK_HIP(hipEventRecord(e0, c.stream));
            for (int k = 0; k < reps; ++k)
                assemble_impl<double>(c, flags, c.K64.p, c.G64.p, c.GT64.p, c.M64.p, c.F64.p,
                                      mg ? c.mg[0].A64.p : nullptr, mg ? c.mg[0].E64.p : nullptr,
                                      mg ? c.RK64.p : nullptr);
            STK_HIP(hipEventRecord(e1, c.stream));
            STK_HIP(hipEventSynchronize(e1));
            float ms = 0;
            STK_HIP(hipEventElapsedTime(&ms, e0, e1));
            *avg_ms = ms / reps;
            // compulsory bytes: per-element inputs, per-node inputs, one write of every operator
            const double in = (double)c.nel * (6 * 4 + 6 * 8 + 6 * 8 + 36 * 8 + 81) + (double)c.nnod * (8 + 16);
            const double out = 8.0 * (4 * c.KN.nstore() + 4 * c.GN.nnz() + c.MN.nnz() + c.nU) +
                                (mg ? 8.0 * (4 * c.mg[0].pat.nnz() + 36 * c.nel) : 0.0);
            *bytes = in + out;
            // the timed launches accumulated into the operators: re-assemble cleanly
            STK_HIP(hipStreamSynchronize(c.stream));
            return stk_assemble(ctx, flags);
        }
        *avg_ms = time_kernel(c, kernel, reps, bytes);
        return (int)STK_OK;
    });
}

void stk_default_train_opts(stk_train_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->snapshots = 2000;
    o->per_snapshot = 512;
    o->epochs = 4;
    o->batch = 4096;
    o->ratio_kind = 1;
    o->reserved = 0;
    o->lr = 3e-3;
    o->lr_final = 1e-4;
    o->seed = 0x5EED2000ull;
}

int stk_set_fabric(stk_ctx *ctx, const double *phi, const double *delta) {
    if (!ctx || !phi || !delta) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        for (int64_t e = 0; e < c.nel; ++e)
            STK_CHECK(std::isfinite(phi[e]) && delta[e] > 0, STK_E_ARG, "bad fabric value");
        c.d_phi.upload(phi, c.nel, c.stream);
        c.d_delta.upload(delta, c.nel, c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));
        c.have_fabric = true;
        return (int)STK_OK;
    });
}

int stk_mlp_train(stk_ctx *ctx, const stk_train_opts *opts, stk_train_report *report) {
    if (!ctx) return STK_E_ARG;
    stk_train_report local{};
    stk_train_report &rep = report ? *report : local;
    std::memset(&rep, 0, sizeof(rep));
    return guarded(ctx, [&]() {
        stk_train_opts o;
        stk_default_train_opts(&o);
        if (opts) o = *opts;
        mlp_train(ctx->c, o, rep);
        return (int)STK_OK;
    });
}

int stk_mlp_params(stk_ctx *ctx, float *params, float *norm, int32_t set) {
    if (!ctx || !params || !norm) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        if (!c.mlp) {
            STK_CHECK(set, STK_E_STATE, "no model");
            c.mlp = new MLP();
            c.mlp->init(1, c.stream);
        }
        MLP &net = *c.mlp;
        if (set) {
            net.w.upload(params, mlp::kNParam, c.stream);
            std::memcpy(net.norm, norm, sizeof(net.norm));
            net.dnorm.upload(net.norm, 22, c.stream);
            net.pack(c.stream);
            net.trained = true;
        } else {
            net.w.download(params, mlp::kNParam, c.stream);
            std::memcpy(norm, net.norm, sizeof(net.norm));
        }
        STK_HIP(hipStreamSynchronize(c.stream));
        return (int)STK_OK;
    });
}

int stk_mlp_forward(stk_ctx *ctx, int64_t n, const float *X, float *Y) {
    if (!ctx || !X || !Y || n <= 0) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.mlp, STK_E_STATE, "no model");
        DBuf<float> x(5 * n), y(6 * n);
        x.upload(X, 5 * n, c.stream);
        c.mlp->forward(x.p, n, y.p, true, c.stream);
        y.download(Y, 6 * n, c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));
        return (int)STK_OK;
    });
}

int stk_mlp_gradient(stk_ctx *ctx, int64_t n, const float *X, const float *T, float *grad, double *loss) {
    if (!ctx || !X || !T || !grad || n <= 0) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.mlp, STK_E_STATE, "no model");
        DBuf<float> x(5 * n), t(6 * n);
        DBuf<double> lb;
        x.upload(X, 5 * n, c.stream);
        t.upload(T, 6 * n, c.stream);
        const double l = c.mlp->gradient(x.p, t.p, n, c.stream, lb);
        if (loss) *loss = l;
        c.mlp->grad.download(grad, mlp::kNParam, c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));
        return (int)STK_OK;
    });
}

int stk_local_schur(stk_ctx *ctx, double *S, float *X) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.have_fields && c.have_fabric, STK_E_STATE, "fields and fabric needed");
        STK_CHECK(c.structured, STK_E_NOMESH, "the learned model needs a structured annulus (Grid_split1 or Grid)");
        DBuf<double> s(6 * c.nel);
        DBuf<float> x(5 * c.nel);
        SchurSampleHost h;
        h.nsamp = c.nel; h.mode = 0; h.X = x.p; h.S = s.p;
        launch_schur_samples(c, h);
        if (S) s.download(S, 6 * c.nel, c.stream);
        if (X) x.download(X, 5 * c.nel, c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));
        return (int)STK_OK;
    });
}

int stk_learned_setup(stk_ctx *ctx) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        learned_setup(ctx->c);
        return (int)STK_OK;
    });
}

int stk_updatevelocity(stk_ctx *ctx, double *U, double *Ucart, double *omega) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.U && c.have_fields, STK_E_STATE, "no solution");
        c.d_Ucart.alloc(c.nU);
        const double w = update_velocity(c, c.d_Ucart.p);
        if (omega) *omega = w;
        if (U) c.U.download(U, c.nU, c.stream);
        if (Ucart) c.d_Ucart.download(Ucart, c.nU, c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));
        return STK_OK;
    });
}

int stk_set_stress_fields(stk_ctx *ctx, const double *eta_ip, int32_t nip, const double *const C[6]) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(eta_ip, STK_E_ARG, "NULL eta.ip");
        stress_set_rule(c, nip);
        stress_prepare(c);
        const int64_t plane = c.nel * (int64_t)nip;
        // Julia nel x ncol column-major: the first nip columns are contiguous
        c.d_seta.upload(eta_ip, plane, c.stream);
        if (C) {
            std::vector<double> buf(6 * plane);
            for (int k = 0; k < 6; ++k) {
                STK_CHECK(C[k], STK_E_ARG, "NULL tensor component");
                std::copy(C[k], C[k] + plane, buf.begin() + k * plane);
            }
            c.d_sC.upload(buf, c.stream);
            STK_HIP(hipStreamSynchronize(c.stream));
            c.s_aniso = true;
        } else {
            c.s_aniso = false;
        }
        STK_HIP(hipStreamSynchronize(c.stream));
        c.have_sfields = true;
        return STK_OK;
    });
}

int stk_stress(stk_ctx *ctx, const double *Ucart, double *const out[8]) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.have_sfields, STK_E_STATE, "stress fields not set");
        DBuf<double> &u = c.staged("stress.u");
        const double *Ud = c.d_Ucart.p;
        if (Ucart) {
            u.upload(Ucart, c.nU, c.stream);
            Ud = u.p;
        }
        STK_CHECK(Ud, STK_E_STATE, "no device velocity: call stk_updatevelocity or pass Ucart");
        const int64_t plane = c.nel * (int64_t)c.s_nip;
        DBuf<double> &o = c.staged("stress.out");
        o.alloc(8 * plane);
        launch_stress(c, Ud, o.p);
        if (out)
            for (int k = 0; k < 8; ++k)
                if (out[k]) STK_HIP(hipMemcpyAsync(out[k], o.p + k * plane, plane * sizeof(double),
                                                   hipMemcpyDeviceToHost, c.stream));
        STK_HIP(hipStreamSynchronize(c.stream));
        return STK_OK;
    });
}

int stk_thermal_solve(stk_ctx *ctx, double *T, const double *rho, double K, double Cp, double dQdT, double dt,
                      const int64_t *fixed_nodes, const double *fixed_values, int64_t nfixed, double rtol,
                      stk_report *report) {
    if (!ctx || !T || !rho || (nfixed > 0 && (!fixed_nodes || !fixed_values)) || nfixed < 0) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        std::vector<int64_t> fx(nfixed);
        for (int64_t k = 0; k < nfixed; ++k) {
            STK_CHECK(fixed_nodes[k] >= 1 && fixed_nodes[k] <= c.nnod, STK_E_SHAPE, "fixed node out of range");
            fx[k] = fixed_nodes[k] - 1;
        }
        DBuf<double> &t = c.thw_io.t, &r = c.thw_io.r, &vf = c.thw_io.vf;
        DBuf<int64_t> &f = c.thw_io.f;
        t.upload(T, c.nnod, c.stream);
        r.upload(rho, c.nnod, c.stream);
        f.upload(fx, c.stream);
        vf.upload(fixed_values, nfixed, c.stream);
        ThermalStep s{K, Cp, dQdT, dt, rtol > 0 ? rtol : 1.4901161193847656e-8, 20000};
        stk_report rep;
        thermal_solve(c, t.p, r.p, s, f.p, vf.p, nfixed, rep);
        t.download(T, c.nnod, c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));
        if (report) *report = rep;
        return rep.converged ? (int)STK_OK : (int)STK_E_NOTCONV;
    });
}

int stk_stats(stk_ctx *ctx, const double *U, const double *T, double out[4]) {
    if (!ctx || !T || !out) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        DBuf<double> &u = c.staged("stats.u"), &t = c.staged("stats.t");
        const double *Ud = c.U.p;
        if (U) {
            u.upload(U, c.nU, c.stream);
            Ud = u.p;
        }
        STK_CHECK(Ud, STK_E_STATE, "no device solution: solve first or pass U");
        t.upload(T, c.nnod, c.stream);
        statistics(c, Ud, t.p, out);
        return STK_OK;
    });
}

int stk_stats_noalloc(stk_ctx *ctx, const double *U, const double *T, double *out) {
    if (!ctx || !U || !T || !out) return STK_E_ARG;
    return guarded(ctx, [&]() {
        statistics_noalloc(ctx->c, U, T, out);
        return STK_OK;
    });
}

int stk_deformation_gradient(stk_ctx *ctx, const double *U, double dt, double *F) {
    if (!ctx || !F) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.s_nip > 0, STK_E_STATE, "stress rule not set (stk_set_stress_fields)");
        stress_prepare(c);
        DBuf<double> &u = c.staged("defgrad.u");
        const double *Ud = c.U.p;
        if (U) {
            u.upload(U, c.nU, c.stream);
            Ud = u.p;
        }
        STK_CHECK(Ud, STK_E_STATE, "no device solution: solve first or pass U");
        const int64_t n = 4 * c.nel * (int64_t)c.s_nip;
        DBuf<double> &f = c.staged("defgrad.f");
        f.upload(F, n, c.stream);
        launch_deformation_gradient(c, Ud, dt, f.p);
        f.download(F, n, c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));
        return STK_OK;
    });
}

int stk_fse(stk_ctx *ctx, double *F, int32_t nip, int32_t mode, double eps, double *fse) {
    if (!ctx || !F) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(nip == 6 || nip == 7, STK_E_SHAPE, "F needs 6 or 7 points per element");
        STK_CHECK(mode >= STK_FSE_NO_HEALING && mode <= STK_FSE_ANNEALING_DESTRUCTION, STK_E_ARG,
                  "unknown healing mode");
        STK_CHECK(mode == STK_FSE_NO_HEALING || eps > 0, STK_E_ARG, "the destruction ratio eps must be positive");
        const int64_t npt = c.nel * (int64_t)nip;
        DBuf<double> &f = c.staged("fse.f");
        f.upload(F, 4 * npt, c.stream);
        c.fse_nip = 0;
        c.d_fse.alloc(6 * npt);
        launch_fse(c, f.p, npt, mode, eps, c.d_fse.p);
        if (mode != STK_FSE_NO_HEALING) f.download(F, 4 * npt, c.stream);
        DBuf<double> &pts = c.staged("fse.points");
        if (fse) {
            pts.alloc(6 * npt);
            fse_layout(c, c.d_fse.p, pts.p, npt, true);
            pts.download(fse, 6 * npt, c.stream);
        }
        STK_HIP(hipStreamSynchronize(c.stream));
        c.fse_nip = nip;
        return STK_OK;
    });
}

int stk_set_dem(stk_ctx *ctx, const double *tensor, const double *axes, int64_t n, double w) {
    if (!ctx || !tensor || !axes) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        c.dem.set = false;
        dem_prepare(c, tensor, axes, n, w);
        return STK_OK;
    });
}

int stk_anisotropic_tensor(stk_ctx *ctx, const double *fse, int32_t nip, const double *shape_a1,
                           const double *shape_r1, const double *shape_r2, int64_t nshape, int32_t install,
                           double *const out[6]) {
    if (!ctx) return STK_E_ARG;
    return guarded(ctx, [&]() {
        Context &c = ctx->c;
        STK_CHECK(c.dem.set, STK_E_STATE, "no DEM table: call stk_set_dem first");
        STK_CHECK(nip == 6 || nip == 7, STK_E_SHAPE, "the ellipses need 6 or 7 points per element");
        STK_CHECK(shape_r1 && shape_r2, STK_E_ARG, "NULL fabric shape");
        STK_CHECK((install & ~(STK_INSTALL_ASSEMBLY | STK_INSTALL_STRESS)) == 0, STK_E_ARG, "unknown install flags");
        STK_CHECK(!(install & STK_INSTALL_ASSEMBLY) || c.have_fields, STK_E_STATE,
                  "install into the assembly needs its fields (stk_set_fields) first");
        STK_CHECK(!(install & STK_INSTALL_STRESS) || c.have_sfields, STK_E_STATE,
                  "install into the stress fields needs stk_set_stress_fields first");
        const int64_t npt = c.nel * (int64_t)nip;
        DBuf<double> &f = c.staged("aniso.fse"), &fp = c.staged("aniso.fse_points"), &sn = c.staged("aniso.a1"),
                     &s1 = c.staged("aniso.r1"), &s2 = c.staged("aniso.r2");
        const double *dfse = c.d_fse.p;
        if (fse) {
            fp.upload(fse, 6 * npt, c.stream);
            f.alloc(6 * npt);
            fse_layout(c, fp.p, f.p, npt, false);
            dfse = f.p;
        } else {
            STK_CHECK(dfse && c.fse_nip == nip, STK_E_STATE,
                      "no device ellipses with this nip: call stk_fse first or pass fse");
        }
        ShapeIn shape;
        if (shape_a1) {
            STK_CHECK(nshape >= 2, STK_E_SHAPE, "the fabric-shape table needs at least two nodes");
            for (int64_t k = 0; k + 1 < nshape; ++k)
                STK_CHECK(shape_a1[k] < shape_a1[k + 1], STK_E_ARG, "the fabric-shape nodes must increase strictly");
            sn.upload(shape_a1, nshape, c.stream);
            s1.upload(shape_r1, nshape, c.stream);
            s2.upload(shape_r2, nshape, c.stream);
            shape.nodes = sn.p;
            shape.nshape = nshape;
        } else {
            STK_CHECK(nshape == npt, STK_E_SHAPE, "per-point fabric shapes need nel x nip values");
            s1.upload(shape_r1, npt, c.stream);
            s2.upload(shape_r2, npt, c.stream);
        }
        shape.r1 = s1.p;
        shape.r2 = s2.p;
        const int64_t plane7 = 7 * c.nel;
        c.d_aniso.alloc(6 * plane7);
        launch_anisotropic_tensor(c, dfse, nip, shape, c.d_aniso.p);
        if (install & STK_INSTALL_ASSEMBLY) {
            // assembly tensor: 6 components x (nel x 6), the first six columns of each plane
            c.d_C.alloc(36 * c.nel);
            for (int k = 0; k < 6; ++k)
                STK_HIP(hipMemcpyAsync(c.d_C.p + k * 6 * c.nel, c.d_aniso.p + k * plane7, 6 * c.nel * sizeof(double),
                                       hipMemcpyDeviceToDevice, c.stream));
            c.have_C = true;
        }
        if (install & STK_INSTALL_STRESS) {
            const int64_t plane = c.nel * (int64_t)c.s_nip;
            c.d_sC.alloc(6 * plane);
            for (int k = 0; k < 6; ++k)
                STK_HIP(hipMemcpyAsync(c.d_sC.p + k * plane, c.d_aniso.p + k * plane7, plane * sizeof(double),
                                       hipMemcpyDeviceToDevice, c.stream));
            c.s_aniso = true;
        }
        if (out)
            for (int k = 0; k < 6; ++k)
                if (out[k]) STK_HIP(hipMemcpyAsync(out[k], c.d_aniso.p + k * plane7, plane7 * sizeof(double),
                                                   hipMemcpyDeviceToHost, c.stream));
        STK_HIP(hipStreamSynchronize(c.stream));
        return STK_OK;
    });
}

}  // extern "C"
