/*
 * dimspers_stokes.h -- C ABI of the MI355X Stokes hot path.
 *
 * The reference has no FFI: its boundary is a set of Julia functions.  Each
 * entry point below names the reference function it stands in for
 * (file:line in the DimsPers tree).  Conventions follow the reference:
 *   - index arrays are int64, 1-based, column-major (Julia layout);
 *   - floating-point inputs are FP64, column-major;
 *   - sparse exports are CSC with 1-based colptr/rowval;
 *   - host pointers are only read/written during the call.
 * Every function returns STK_OK (0) or a negative STK_E_* code; the message
 * is available from stk_last_error(ctx) (or stk_last_error(NULL) when
 * stk_create itself fails).  A context owns device memory and one HIP
 * stream on the device that was current at stk_create; it is not thread-safe.
 */
#ifndef DIMSPERS_STOKES_H
#define DIMSPERS_STOKES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct stk_ctx stk_ctx;

enum {
    STK_OK = 0,
    STK_E_HIP = -1,       /* HIP runtime error (message carries hipGetErrorString)     */
    STK_E_SHAPE = -2,     /* inconsistent sizes / indices                              */
    STK_E_STATE = -3,     /* call out of order (e.g. solve before assemble)            */
    STK_E_NOTCONV = -4,   /* solver stopped at its iteration cap                       */
    STK_E_NAN = -5,       /* non-finite value met in the solve                         */
    STK_E_ARG = -6,       /* bad option value                                          */
    STK_E_NOMESH = -7     /* mesh is not the structured annulus the multigrid needs   */
};

/* Assembly flags (stk_assemble). */
enum {
    STK_ASM_FP64 = 1,        /* compute in FP64, write FP64 operators                    */
    STK_ASM_FP32 = 2,        /* FP32 operators for the inner solves (with FP64: rounded) */
    STK_ASM_CARTESIAN = 4,   /* skip the TT^T . TT rotation (assembly_stokes_cylindric)  */
    STK_ASM_NO_BC = 8,       /* keep rows/columns of fixed dofs (no _apply_bcs)          */
    STK_ASM_ISOTROPIC = 16,  /* isotropic D (StokesSolver.jl:688-742), C6 ignored       */
    STK_ASM_NO_MG = 32,      /* do not build the multigrid hierarchy                     */
    STK_ASM_FP16_SMOOTHER = 64,/* fine-level smoother on an FP16 copy of S^-1 K S^-1       */
    STK_ASM_7POINT = 128,    /* 7-point rule (ip_triangle7) of the threaded assembler,   */
                             /* AssemblerStokes.jl:174 (needs 7 columns in stk_set_fields) */
    STK_ASM_CLAMP = 256      /* clamp every eta * C entry to [1e18, 1e25] as the top-level */
                             /* _fill_Kblk! does for dimensional runs (StokesSolver.jl:756) */
};

/* Which operator stk_export_csc / stk_spmv address. */
enum {
    STK_MAT_K = 0,    /* velocity block, nU x nU                 */
    STK_MAT_G = 1,    /* gradient, nU x nP                       */
    STK_MAT_GT = 2,   /* divergence G^T, nP x nU                 */
    STK_MAT_M = 3,    /* 1/eta pressure mass, nP x nP            */
    STK_MAT_S = 4,    /* learned/exact local-Schur sum, nP x nP  */
    STK_MAT_KT = 5,   /* thermal diffusion matrix, nnod x nnod   */
    STK_MAT_MT = 6,   /* thermal mass matrix rho Cp, nnod x nnod */
    STK_MAT_R = 7,    /* multigrid restriction P2 -> P1 (rotated */
                      /* frames, fixed dofs masked), 2 nP x nU   */
    STK_MAT_RK = 8,   /* R K (K with boundary conditions), the   */
                      /* coarse residual operator, 2 nP x nU     */
    STK_MAT_LEVEL = 16,   /* + l: the operator of multigrid level l  */
                          /* (l = 0: P1 on the mesh's vertices), */
                          /* 2 nvert_l x 2 nvert_l                   */
    STK_MAT_PROLONG = 48, /* + l: the prolongation into the next */
                          /* finer space from level l (l = 0: into */
                          /* the P2 velocity, = R^T), n_finer x 2 nvert_l */
    STK_MAT_LEVEL_RA = 80, /* + l (l >= 1): R A, the restriction into */
                          /* level l times the operator of level l-1, */
                          /* the coarse residual operator of the V-cycle */
                          /* (after a solver set-up), 2 nvert_l x 2 nvert_l-1 */
    STK_MAT_COARSE_INV = 112 /* the dense inverse of the coarsest level's */
                          /* operator (after a solver set-up; FP64 as     */
                          /* computed), every entry stored, n x n         */
};

/* Fine-level smoother of the FP32 multigrid V-cycle (stk_opts.fine_smoother). */
enum {
    STK_SMOOTH_GS = 0,      /* multicolour block Gauss-Seidel: forward sweep before the coarse */
                            /* correction, backward sweep after (default)                     */
    STK_SMOOTH_CHEB = 1     /* Chebyshev of degree cheb_degree with point Jacobi               */
};

/* Solver methods (stk_opts.method). */
enum {
    STK_SOLVE_PLAIN = 0,    /* FP64 Schur PCG, lumped 1/eta-mass preconditioner (reference) */
    STK_SOLVE_MIXED = 1,    /* FP64 refinement around FP32 Schur PCG                        */
    STK_SOLVE_LEARNED = 2   /* MIXED with the learned local-Schur preconditioner            */
};

/* Inner velocity solver (stk_opts.inner). */
enum {
    STK_INNER_MG = 0,       /* CG + geometric multigrid V-cycle   */
    STK_INNER_JACOBI = 1    /* CG + point Jacobi (plain method only: */
                            /* the mixed and learned methods need */
                            /* the multigrid, STK_E_ARG otherwise)  */
};

/* Treatment of the discrete rigid rotation (stk_opts.null_mode).  With free slip
 * on both walls the lowest eigenvector v of K is a slightly compressible rigid
 * rotation whose eigenvalue falls like h^6 (about 1e-16 of the largest at 64 x 512
 * cells), so K is numerically singular. */
enum {
    STK_NULL_PROJECT = 0,   /* velocities orthogonal to v (well posed at every size)        */
    STK_NULL_DEFLATE = 1    /* also add (v.y / lambda) v to every inner solve: the exact    */
                            /* K^-1 of lu(K[ifree, ifree]) where FP64 still resolves lambda */
};

/* Stopping rule (stk_opts.stop). */
enum {
    STK_STOP_SCHUR = 0,     /* ||r_k|| < rtol ||r_0|| on the Schur residual (IterativeSolvers.jl:118,156) */
    STK_STOP_FULL = 1       /* ||[f + G p - K u ; -G^T u]|| < rtol ||[f ; 0]|| in FP64, with the  */
                            /* constant removed from -G^T u and, for STK_NULL_PROJECT, the v      */
                            /* component removed from f + G p - K u                              */
};

typedef struct stk_opts {
    int32_t method;         /* STK_SOLVE_*                                         */
    int32_t inner;          /* STK_INNER_*                                         */
    int32_t stop;           /* STK_STOP_*                                          */
    int32_t itmax;          /* outer (Schur) iteration cap; reference: 200         */
    int32_t itmin;          /* minimum outer iterations; reference: 10             */
    int32_t inner_itmax;    /* inner CG cap                                        */
    int32_t refine_max;     /* mixed/learned: FP64 refinement sweeps cap           */
    int32_t cheb_degree;    /* multigrid Chebyshev smoothing degree                */
    int32_t null_mode;      /* STK_NULL_*                                          */
    int32_t pc_sweeps;      /* learned: damped Jacobi sweeps on the learned Schur  */
                            /* operator per preconditioner application (0 = 3)     */
    double rtol;            /* outer tolerance; reference: 1e-8                    */
    double inner_rtol;      /* inner CG relative tolerance (plain method)          */
    double mixed_inner_rtol;/* inner CG tolerance inside the FP32 Schur PCG         */
    double mixed_schur_rtol;/* full-residual reduction asked of one FP32 correction */
    double stall_ratio;     /* learned: fall back when a sweep's Schur PCG misses its */
                            /* target and the full residual falls by less than this */
    int32_t inner_warm;     /* 1: each inner solve of the Schur PCG starts from      */
                            /* beta * z of the previous one (q_k = d_k + beta q_k-1) */
    int32_t fine_smoother;  /* STK_SMOOTH_*: smoother of the fine multigrid level  */
                            /* in the FP32 inner solves                            */
    int32_t warm_start;     /* 0: U and P start at zero (StokesPcCG, :101-102);    */
                            /* 1: start from the context's solution (the last      */
                            /* solve's, or stk_set_solution's): the pressure is    */
                            /* kept as in StokesPcCG_pardiso (IterativeSolvers.jl: */
                            /* 7-17, Rhs .+= GG*P), the first pass solves for the  */
                            /* correction from the full residual of (U, P)         */
    int32_t pressure_pc;    /* STK_PPC_*: pressure preconditioner of the plain     */
                            /* method (_preconditioner, :189-217)                  */
} stk_opts;

/* Pressure preconditioner of the plain FP64 Schur PCG (stk_opts.pressure_pc). */
enum {
    STK_PPC_LUMPED = 0,     /* inverse row sums of MM (:jacobi = lumpedpreconditioner, StokesPcCG) */
    STK_PPC_IC = 1          /* incomplete Cholesky of MM (:incomplete_chol = lldl, the            */
                            /* preconditioner of StokesPcCG_pardiso, :32, :242-244)               */
};

typedef struct stk_report {
    int32_t converged;          /* 1 if the stopping rule was met               */
    int32_t outer_iterations;   /* Schur PCG iterations (all sweeps)            */
    int64_t inner_iterations;   /* inner CG iterations (all)                    */
    int32_t refinements;        /* FP64 refinement sweeps (mixed / learned)     */
    int32_t fallbacks;          /* learned -> plain FP64 fallbacks              */
    double schur_residual;      /* final ||r_k|| / ||r_0|| of the last Schur PCG */
    double full_residual;       /* FP64 full-system relative residual           */
    double solve_ms;            /* wall time of the solve on the stream         */
    double bytes;               /* algorithmic bytes of the solve's kernels: sparse and dense */
                                /* products and Gauss-Seidel sweeps by their operands (blocks, */
                                /* indices, slice pointers, the gathered vector once, the     */
                                /* vectors of the epilogue), element-wise kernels by the      */
                                /* vectors they read and write; `launches_floor` of them state */
                                /* no count and are taken at 8 bytes per element              */
    int64_t launches;           /* kernel launches of the solve                   */
    int64_t launches_floor;     /* of which counted at the 8-byte floor           */
} stk_report;

/* Fill *opts with the defaults of the reference StokesPcCG. */
void stk_default_opts(stk_opts *opts);

/* Mesh set-up.  Replaces sparsitystokes (src/Algebra/Sparsity.jl:10-29),
 * rotation_matrix (src/Stokes/StokesSolver.jl:203-245) and the element
 * coordinates of MAIN.jl:62-66; fixed_dofs is UBC.Omega
 * (src/Setup/boundary_conditions.jl:15-37), 1-based.
 * e2n: 6 x nel (P2, local order v1 v2 v3 m12 m23 m31); el_theta/el_r: 3 x nel
 * vertex polar coordinates after fixangles!; node_theta/node_r: gr.theta, gr.r. */
int stk_create(stk_ctx **ctx, int64_t nel, int64_t nnod, int64_t nvert,
               const int64_t *e2n, const double *el_theta, const double *el_r,
               const double *node_theta, const double *node_r,
               const int64_t *fixed_dofs, int64_t nfixed);

void stk_destroy(stk_ctx *ctx);

/* Message of the last failing call on ctx (NULL: of the last failing stk_create).  A call that
 * succeeds clears it; the pointer stays valid until the next call on ctx, so copy the text
 * before calling again. */
const char *stk_last_error(const stk_ctx *ctx);

/* Sizes: out[0..6] = nel, nnod, nvert, nU, nP, colours, multigrid levels. */
int stk_info(const stk_ctx *ctx, int64_t *out7);

/* Which colouring stk_colouring returns. */
enum {
    STK_COLOUR_ELEMENTS = 0,  /* element colours of the assembly scatter (nel values)          */
    STK_COLOUR_NODES = 1      /* node colours of the multicolour Gauss-Seidel smoother (nnod)  */
};

/* The colourings built by stk_create, for checking: no two elements of one colour share a
 * node (the conflict-free scatter of color_mesh, src/Grid/coloring.jl:5-64, whose colouring
 * check_colored_mesh :97-111 is the reference's counterpart), and no two coupled nodes share a
 * colour (the smoother's independent sets).  colour receives 0-based colours, *ncolour their
 * number. */
int stk_colouring(stk_ctx *ctx, int32_t which, int32_t *colour, int32_t *ncolour);

/* Upload the per-step fields assembly_stokes_cylindric reads
 * (StokesSolver.jl:247-249): g, rho.rho (nnod), eta.ip (nel x nip column-major,
 * nip >= 6; the 6-point rule uses the first 6 columns, :608, STK_ASM_7POINT the first 7)
 * and the six StiffnessTensor arrays (src/Rheology/anisotropy.jl:1-8) in the order
 * eta11, eta33, eta55, eta13, eta15, eta35, each nel x nip.  C may be NULL for the
 * isotropic law. */
int stk_set_fields(stk_ctx *ctx, double g, const double *rho_node,
                   const double *eta_ip, int64_t nip, const double *const C[6]);

/* Element assembly on the GPU.  Replaces assembly_stokes_cylindric
 * (StokesSolver.jl:247-404) and, unless STK_ASM_CARTESIAN / STK_ASM_NO_BC,
 * _prepare_matrices (:192-201) and _apply_bcs (:186-190).  flags: STK_ASM_*. */
int stk_assemble(stk_ctx *ctx, int32_t flags);

/* Export an assembled operator as 1-based CSC (FP64).  Call with
 * colptr/rowval/nzval NULL to get *nnz (and *nrows, *ncols) first.
 * lower_only = 1 gives the lower triangle (KK, MM as the reference returns them). */
int stk_export_csc(stk_ctx *ctx, int32_t which, int32_t lower_only,
                   int64_t *nrows, int64_t *ncols, int64_t *nnz,
                   int64_t *colptr, int64_t *rowval, double *nzval);

/* Export the assembled right-hand side (Fb, nU), FP64. */
int stk_export_rhs(stk_ctx *ctx, double *F);

/* y = A x with an assembled operator (parity checks of the SpMV kernels).
 * precision 64 or 32 selects the FP64 or FP32 copy; x, y are host FP64. */
int stk_spmv(stk_ctx *ctx, int32_t which, int32_t precision, const double *x, double *y);

/* Schur-complement PCG.  Replaces StokesPcCG (src/IterativeSolvers.jl:94-168),
 * its _CholeskyFactorizationSolve inner solve (src/PardisoSolvers.jl:247-278)
 * and StokesPcCG_gpu (src/src/GPUsolver.jl:2-86).  U and P start at zero as
 * in the reference (:101-102).  Returns STK_E_NOTCONV when the cap is hit
 * (the reference returns silently); the report is filled either way. */
int stk_solve(stk_ctx *ctx, const stk_opts *opts, stk_report *report);

/* Copy the last solution to the host (U: nU, P: nP; polar dof frame). */
int stk_get_solution(stk_ctx *ctx, double *U, double *P);

/* Set the solution a solve with stk_opts.warm_start = 1 starts from: the incoming U and P of
 * StokesPcCG_pardiso (src/IterativeSolvers.jl:1-17), which keeps P and solves for U.  U (nU) and
 * P (nP) are host FP64 in the polar dof frame; either may be NULL (zero). */
int stk_set_solution(stk_ctx *ctx, const double *U, const double *P);

/* StokesPcCG's operators from the caller instead of stk_assemble: the matrices StokesPcCG(U, P, KK,
 * MM, GG, Rhs, ifree) receives (src/IterativeSolvers.jl:94; solveStokes passes them at
 * src/Stokes/StokesSolver.jl:41-46, the threaded assembler at src/Stokes/AssemblerStokes.jl:62).
 * K: the rotated velocity block of _prepare_matrices (nU x nU, full symmetric, or its lower
 * triangle with K_lower = 1); MM: the pressure mass matrix (nP x nP) as assembly returns it, lower
 * triangle with M_lower = 1 (_preconditioner symmetrises it, :190); GG: nU x nP; all 1-based CSC
 * (colptr of length ncols + 1, rowval, nzval), with every entry inside the mesh's P2-P1 pattern
 * (STK_E_SHAPE otherwise; explicit zeros are allowed).  Rhs (nU) is the right-hand side after
 * _apply_bcs.  The fixed dofs are those given to stk_create (UBC.Omega, the complement of ifree):
 * their rows and columns of K and rows of GG are dropped, as the inner solve on K[ifree, ifree]
 * (src/PardisoSolvers.jl:247-256) and the free-dof solution leave them out.  flags: STK_ASM_FP64
 * and/or STK_ASM_FP32 (rounded copies for the mixed method).  Then stk_solve runs on these
 * operators; the multigrid, which is built from the element matrices of stk_assemble, is not
 * available (STK_INNER_JACOBI, plain method). */
int stk_set_operators(stk_ctx *ctx, int32_t K_lower, const int64_t *K_colptr, const int64_t *K_rowval,
                      const double *K_nzval, int32_t M_lower, const int64_t *M_colptr, const int64_t *M_rowval,
                      const double *M_nzval, const int64_t *G_colptr, const int64_t *G_rowval,
                      const double *G_nzval, const double *Rhs, int32_t flags);

/* Net-rotation removal and Cartesian velocity.  Replaces updatevelocity2
 * (StokesSolver.jl:170-184) with volume_integral (src/Stress/FSE.jl:412-464).
 * Operates on the device solution and keeps the Cartesian velocity on the device
 * (the default input of stk_stress); U (polar, nU) and Ucart (nU) may be NULL. */
int stk_updatevelocity(stk_ctx *ctx, double *U, double *Ucart, double *omega);

/* Fields of the stress post-processor: eta.ip and, for the anisotropic law, the six
 * StiffnessTensor arrays (order of stk_set_fields), each nel x ncol column-major with
 * ncol >= nip; the first nip columns are the points of the nip-point rule (6 or 7;
 * MAIN.jl:164 builds the stress shape functions with 7).  C = NULL selects the isotropic
 * method of _stress! (src/Stress/StressCalculation.jl:23-127), else the anisotropic one
 * (:129-242). */
int stk_set_stress_fields(stk_ctx *ctx, const double *eta_ip, int32_t nip, const double *const C[6]);

/* stress! (src/Stress/StressCalculation.jl:244-249): strain rate, its second invariant,
 * deviatoric stress and its second invariant at every point of the stress rule, from the
 * Cartesian velocity Ucart (x, z per node, nU; MAIN.jl:241-243 passes Ucart).  Ucart = NULL
 * uses the device velocity of the last stk_updatevelocity.  out[8] receive eps.xx, eps.zz,
 * eps.xz, epsII, tau.xx, tau.zz, tau.xz, tauII, each nel x nip column-major; any may be NULL. */
int stk_stress(stk_ctx *ctx, const double *Ucart, double *const out[8]);

/* deformation_gradient! (src/Stress/StressCalculation.jl:251-349): one RK4 step of
 * F' = dt (grad U)^T F at every point of the stress rule.  F (inout) is the Julia
 * Matrix{SMatrix{2,2}} nel x nip: 4 doubles (column-major 2x2) per point, points column-
 * major.  U = NULL uses the device solution in the node frames, which is what MAIN.jl:254
 * passes (the velocity after stk_updatevelocity's rotation removal). */
int stk_deformation_gradient(stk_ctx *ctx, const double *U, double dt, double *F);

/* write_stats (src/IO/IOmanager.jl:277-322) without the file output: out[0] = Urms =
 * sqrt(int |U|^2 / area), out[1] = <T> = int T / area (volume_integral_cartesian,
 * src/Stress/FSE.jl:466-514: straight triangles, 6-point rule, P2 interpolation of the nodal
 * values), out[2], out[3] = Nusselt number at the top and bottom walls from the nodes one
 * radius level inside them (ScratchNusselt, IOmanager.jl:338-354).  U: velocity, 2 per node
 * (NULL = the device solution); T: temperature at every node (nnod, host). */
int stk_stats(stk_ctx *ctx, const double *U, const double *T, double out[4]);

/* stk_stats that allocates nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight time-step loops, harnesses that forbid
 * hipMalloc).  U (nU: 2 per node), T (nnod) and out (4 doubles, receiving what stk_stats returns) are
 * DEVICE pointers on the context's device; the work is enqueued on the context stream and the call
 * returns without waiting.  Its reduction partials and sums live in one fixed device array of
 * 2 * 1024 + 8 doubles (16 KiB) per device, held until the process ends and shared by every call on
 * that device: calls that overlap on one device (other streams, other threads) race.  The scratch need
 * is bounded whatever the mesh size (grid-stride reductions over at most 1024 blocks); a larger need is
 * refused with STK_E_SHAPE.  STK_E_STATE if the mesh lacks the radius levels write_stats needs. */
int stk_stats_noalloc(stk_ctx *ctx, const double *U, const double *T, double *out);

/* One thermal diffusion step: solveDiffusion_threaded (src/ThermalDiffusion/
 * AssemblerThermal.jl:1-62) with assemble_element! (:117-209; 7-point rule,
 * MAIN.jl:167-169), _prepare_matrices! (ThermalSolver.jl:54-63: Rhs = M T0 + F,
 * A = M + dt K) and _apply_bcs! (:38-41).  T (nnod) holds T0 on entry and the new
 * temperature on return; rho: nnod (the vertex values are used); K, Cp, dQdT: the
 * ThermalParameters fields (src/Setup/Setup.jl:1-12); fixed_nodes: 1-based node ids of
 * TBC.Omega with fixed_values (heated_bottom, boundary_conditions.jl:77-93).  The free
 * block is solved to |r| <= rtol |Rhs_free| (rtol <= 0: 1.5e-8, the default stop of the
 * reference's Krylov.gmres); report: converged, iterations (outer_iterations), relative
 * residual (schur_residual), solve_ms.  K_T and M_T stay available to stk_export_csc. */
int stk_thermal_solve(stk_ctx *ctx, double *T, const double *rho, double K, double Cp, double dQdT, double dt,
                      const int64_t *fixed_nodes, const double *fixed_values, int64_t nfixed, double rtol,
                      stk_report *report);

/* ---------------------------------------------------------------------------
 * Fabric: finite strain ellipses and the DEM viscous tensor.
 * ------------------------------------------------------------------------- */

/* Healing type of FiniteStrain (src/Stress/FSE.jl:39-67) that getFSE dispatches on. */
enum {
    STK_FSE_NO_HEALING = 0,            /* getFSE{NoHealing} / _FSE (:131-161)                    */
    STK_FSE_FABRIC_DESTRUCTION = 1,    /* getFSE{FabricDestructionOnly} / healing_FSE! (:191-231) */
    STK_FSE_ANNEALING_DESTRUCTION = 2  /* getFSE{AnnealingFabricDestruction} (:297-355)          */
};

/* getFSE (src/Stress/FSE.jl:131-231, 297-355): the finite strain ellipse of F F^T at every
 * point of F.  F (inout) has the layout of stk_deformation_gradient (nel x nip points, 4
 * doubles per point, column-major 2x2), nip = 6 or 7 (MAIN.jl:114 uses 7).  With the
 * destruction modes, a point whose a1/a2 >= eps (MAIN.jl:117: 1e3) gets F = I and the unit
 * ellipse; STK_FSE_ANNEALING_DESTRUCTION first divides every F by max |F11| when that exceeds
 * 1e30 (normalize_F!).  fse (may be NULL) receives the FiniteStrainEllipsoid fields (x1, x2,
 * y1, y2, a1, a2) of each point, 6 doubles per point in F's point order.  The ellipses also
 * stay on the device for stk_anisotropic_tensor. */
int stk_fse(stk_ctx *ctx, double *F, int32_t nip, int32_t mode, double eps, double *fse);

/* getDEM (src/Rheology/anisotropy.jl:60-93) from the arrays DEMloader (:48-54) reads out of
 * the HDF5 file: tensor = dataset "C" (n x 9: eta11 eta22 eta33 eta44 eta55 eta66 eta12 eta13
 * eta23), axes = dataset "ax" (n x 3), both column-major.  The table is n / sblk blocks of sblk
 * entries with equal ax[:,1]; w is the viscosity cut-off weakening(dη, φ) of Fabrics.jl, an
 * external package whose value the caller supplies. */
int stk_set_dem(stk_ctx *ctx, const double *tensor, const double *axes, int64_t n, double w);

/* Where stk_anisotropic_tensor also writes its result on the device. */
enum {
    STK_INSTALL_ASSEMBLY = 1,  /* the tensor of stk_set_fields (columns 1-6); assemble again after */
    STK_INSTALL_STRESS = 2     /* the tensor of stk_set_stress_fields (its nip columns)            */
};

/* anisotropic_tensor (src/Rheology/anisotropy.jl:140-222) with get_tensor_and_rotate!
 * (:244-285): per point the DEM entry nearest to the fabric shape, its tensor rotated into the
 * ellipse axis (x1, y1), eta55 clamped to w; the 7th column is the mean of the first six.
 * fse: 6 doubles per point of nel x nip points (nip 6 or 7), or NULL for the device ellipses of
 * the last stk_fse (same nip).  The fabric shape (r1, r2) = fabric_shape(a1, 1,
 * parameterization) of Fabrics.jl is an input: with shape_a1 non-NULL, r1 and r2 sampled at
 * nshape >= 2 strictly increasing a1 nodes, linear between them and constant beyond; with
 * shape_a1 NULL, r1 and r2 at every point (nshape = nel x nip).  out[6] (any may be NULL)
 * receive eta11, eta33, eta55, eta13, eta15, eta35, each nel x 7 column-major (the
 * StiffnessTensor the reference returns).  install: STK_INSTALL_* flags. */
int stk_anisotropic_tensor(stk_ctx *ctx, const double *fse, int32_t nip, const double *shape_a1,
                           const double *shape_r1, const double *shape_r2, int64_t nshape, int32_t install,
                           double *const out[6]);

/* ---------------------------------------------------------------------------
 * Learned local-Schur preconditioner (no counterpart in the reference).  A
 * per-element MLP (5 -> 64 -> 64 -> 64 -> 6, GELU, BF16 weights, FP32
 * accumulation on MFMA) maps (mean log eta, cos 2(phi - theta_c),
 * sin 2(phi - theta_c), log delta, area / mean area of the element's cell) to the log-Cholesky factor
 * of the element's local Schur block S_e = G_e^T (K_e + mu N N^T)^-1 G_e
 * divided by area / eta; the assembled sum of the predicted blocks is the
 * pressure preconditioner of STK_SOLVE_LEARNED.
 * ------------------------------------------------------------------------- */
typedef struct stk_train_opts {
    int32_t snapshots;      /* seeded snapshots of exact local Schur blocks (2000)      */
    int32_t per_snapshot;   /* elements drawn per snapshot                              */
    int32_t epochs;
    int32_t batch;          /* minibatch size (multiple of 32)                          */
    int32_t ratio_kind;     /* delta law of the draws: 0 = 10, 1 = U[1,100], 2 = logU[1,1e3] */
    int32_t reserved;
    double lr, lr_final;    /* Adam step, cosine-decayed to lr_final                    */
    uint64_t seed;
} stk_train_opts;

typedef struct stk_train_report {
    double initial_loss, final_loss, validation_loss;   /* MSE in normalised output units */
    double ms;
    int64_t samples;
    int32_t steps;
    int32_t reserved;
} stk_train_report;

void stk_default_train_opts(stk_train_opts *opts);

/* Per-element fabric director angle phi (radians, Cartesian x-z frame) and
 * normal/shear viscosity ratio delta: the inputs the reference's fabric model
 * turns into the viscous tensor (src/Rheology/anisotropy.jl:140-285). */
int stk_set_fabric(stk_ctx *ctx, const double *phi, const double *delta);

/* Generate the training snapshots from the current mesh and fields with the
 * plain FP64 element matrices and train the MLP on the GPU. */
int stk_mlp_train(stk_ctx *ctx, const stk_train_opts *opts, stk_train_report *report);

/* Read (set = 0) or write (set = 1) the FP32 master parameters (9094 floats:
 * W1 5x64, b1, W2 64x64, b2, W3 64x64, b3, W4 64x6, b4; weights [in][out]) and
 * the normalisation (22 floats: feature means, stds, output means, stds). */
int stk_mlp_params(stk_ctx *ctx, float *params, float *norm, int32_t set);

/* Forward pass on host features (n x 5) -> denormalised outputs (n x 6). */
int stk_mlp_forward(stk_ctx *ctx, int64_t n, const float *X, float *Y);

/* Loss (mean squared error of the normalised outputs) and its gradient with
 * respect to the 9094 parameters for host features / raw targets (n x 6). */
int stk_mlp_gradient(stk_ctx *ctx, int64_t n, const float *X, const float *T, float *grad, double *loss);

/* Exact local Schur blocks of every element for the current fields and fabric:
 * S (nel x 6: S00 S10 S11 S20 S21 S22 in canonical vertex order, divided by
 * area / eta) and the model features X (nel x 5).  Either may be NULL. */
int stk_local_schur(stk_ctx *ctx, double *S, float *X);

/* Run the model on every element and assemble the learned Schur operator used
 * by STK_SOLVE_LEARNED (after stk_assemble). */
int stk_learned_setup(stk_ctx *ctx);

/* Kernels stk_time_kernel can time (on the assembled operators). */
enum {
    STK_KERNEL_SPMV_K32 = 0,    /* y = K x, FP32 BSR 2x2                                  */
    STK_KERNEL_SPMV_K64 = 1,    /* y = K x, FP64 BSR 2x2                                  */
    STK_KERNEL_CHEB_K32 = 2,    /* fused Chebyshev step x' = x + c1 d + c2 D^-1 (b - K x) */
    STK_KERNEL_CHEB_K64 = 3,
    STK_KERNEL_ASSEMBLY = 4,    /* one full element assembly (all colours), FP64          */
    STK_KERNEL_CHEB_KH16 = 5,   /* Chebyshev step on the FP16 scaled copy of K (smoother) */
    STK_KERNEL_COPY = 6,        /* plain copy of 256 MiB (FP32): traffic-counter reference */
    STK_KERNEL_VCYCLE32 = 7,    /* one FP32 multigrid V-cycle (the inner preconditioner)  */
    STK_KERNEL_INNER_IT32 = 8,  /* one FP32 inner CG iteration (V-cycle, K p, updates)    */
    STK_KERNEL_STRESS = 9,      /* stress! on the stress fields (synthetic velocity)      */
    STK_KERNEL_DEFGRAD = 10,    /* deformation_gradient! RK4 step (synthetic velocity)    */
    STK_KERNEL_KP32 = 11,       /* the inner CG product Ap = K p, FP32, with (p . Ap,     */
                                /* v . Ap) block partials                                 */
    STK_KERNEL_GS_BWD32 = 12,   /* backward Gauss-Seidel sweep of the V-cycle with its    */
                                /* dots; time and bytes per colour launch                 */
    STK_KERNEL_FSE = 13,        /* getFSE (fabric destruction) on nel x 7 synthetic F     */
    STK_KERNEL_ANISO = 14,      /* anisotropic_tensor on the device ellipses of the last  */
                                /* stk_fse and the DEM table (16-node fabric-shape table) */
    STK_KERNEL_G32 = 15,        /* y = G p (pressure to velocity), FP32                   */
    STK_KERNEL_GT32 = 16,       /* y = G^T u (velocity to pressure), FP32                 */
    STK_KERNEL_COARSE32 = 17,   /* dense coarsest-level solve (inverse times b), FP32     */
    STK_KERNEL_CGUPD32 = 18,    /* inner CG update z += a p, r -= a P Ap with |r|^2 (the  */
                                /* map pass and the final reduction), FP32                */
    STK_KERNEL_RESTRICT32 = 19  /* level-0 coarse right-hand side R b - (R K) x in one     */
                                /* pass (the V-cycle's restriction), FP32                 */
};

/* Measurement hook with no counterpart in the reference: runs one of the hot
 * kernels `reps` times on the context stream between HIP events and returns
 * the mean duration of one launch (ms) and its algorithmic HBM bytes (values,
 * indices and vectors each moved once).  The operands are the context's own
 * assembled operators and solver work vectors, at their real sizes. */
int stk_time_kernel(stk_ctx *ctx, int32_t kernel, int32_t reps, double *avg_ms, double *bytes);

#ifdef __cplusplus
}
#endif

#endif /* DIMSPERS_STOKES_H */
