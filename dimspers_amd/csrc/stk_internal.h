// Internal declarations shared by the host set-up, the kernels and the solvers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dimspers_stokes.h"

namespace stk {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define STK_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::stk::Error(STK_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define STK_CHECK(cond, code, msg)                           \
    do {                                                     \
        if (!(cond)) throw ::stk::Error((code), (msg));      \
    } while (0)

// Two HIP events destroyed on scope exit, also when a call between them throws.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        STK_HIP(hipEventCreate(&a));
        const hipError_t e = hipEventCreate(&b);
        if (e != hipSuccess) {
            (void)hipEventDestroy(a);
            throw Error(STK_E_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    ~EventPair() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    EventPair(const EventPair &) = delete;
    EventPair &operator=(const EventPair &) = delete;
};

// Owning device buffer.  release() empties it (p null, n zero: its contents are gone) but keeps the
// allocation, which the next alloc() of the same size takes back instead of a hipFree / hipMalloc
// pair (each re-assembly releases and re-allocates the same operators).  Only this buffer reuses
// it, and its new contents are written in stream order after the work that read the old ones.
// purge() and the destructor free the memory.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), spare(o.spare), spare_n(o.spare_n) {
        o.p = o.spare = nullptr;
        o.n = o.spare_n = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            purge();
            p = o.p; n = o.n; spare = o.spare; spare_n = o.spare_n;
            o.p = o.spare = nullptr;
            o.n = o.spare_n = 0;
        }
        return *this;
    }
    ~DBuf() { purge(); }
    void alloc(size_t count) {
        if (count == n && p) return;
        if (!p && count && count == spare_n) {
            p = spare; n = count;
            spare = nullptr; spare_n = 0;
            return;
        }
        purge();
        n = count;
        if (count) STK_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    void release() {
        if (p) {
            if (spare) (void)hipFree(spare);
            spare = p; spare_n = n;
        }
        p = nullptr;
        n = 0;
    }
    void purge() {
        if (p) (void)hipFree(p);
        if (spare) (void)hipFree(spare);
        p = spare = nullptr;
        n = spare_n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) STK_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &h, hipStream_t s) { upload(h.data(), h.size(), s); }
    void download(T *h, size_t count, hipStream_t s) const {
        if (count) STK_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    void zero(hipStream_t s);
    explicit operator bool() const { return p != nullptr; }

  private:
    T *spare = nullptr;   // the allocation release() kept
    size_t spare_n = 0;
};

// Node-level sparsity pattern: row i lists the sorted column nodes of its blocks.
// Optionally the values live in a sliced-ELL order: rows are grouped in slices of
// kSlice consecutive rows, each slice padded to its longest row, and block k of row
// n sits at sptr[n / kSlice] + k * kSlice + n % kSlice (padding: the row's last column, zero
// values).  One lane per row then reads consecutive blocks with its neighbours.
struct Pattern {
    static constexpr int kSlice = 64;
    int64_t nrows = 0, ncols = 0;
    std::vector<int32_t> ptr, col;   // host copies (CSR order)
    DBuf<int32_t> dptr, dcol;        // device copies
    bool sell = false;
    std::vector<int64_t> sptr;       // slice block offsets (nslices + 1)
    std::vector<int32_t> scol;       // columns in sliced order, padded
    DBuf<int64_t> dsptr;
    DBuf<int32_t> dscol;
    int64_t nnz() const { return (int64_t)col.size(); }
    // number of stored blocks (value array length / block size)
    int64_t nstore() const { return sell ? sptr.back() : nnz(); }
    int64_t pos(int64_t row, int64_t k) const {
        return sell ? sptr[row / kSlice] + k * kSlice + row % kSlice : ptr[row] + k;
    }
};

// Work vectors of one multigrid level (or of the inner CG).
template <typename T>
struct MGLevelBufs {
    DBuf<T> b, x, xa, r, d;
};

template <typename T>
struct SchurBufs {
    DBuf<T> v;   // y, z, du (nU) then r, d, q, Sq, rl, pc, dp (nP)
};

// One multigrid level (P1 vertex lattice of the structured annulus, 2 dofs per vertex).  The
// cells are split in two triangles (Grid_split1) or in four around a centre vertex (Grid);
// corner vertex (i, j) is i + (ncr + 1) j, the centre of cell q = i + ncr j is (ncr + 1) nct + q.
struct MGLevel {
    int32_t ncr = 0, nct = 0;        // cells radial / azimuthal
    int32_t split = 2;               // triangles per cell: 2 or 4
    int64_t nvert = 0, nel = 0;
    Pattern pat;                     // vertex -> vertex (BSR 2x2)
    std::vector<int32_t> h_e2v;      // 3 x nel vertex ids (element-major per local vertex)
    DBuf<int32_t> e2v;               // 3 x nel, SoA
    DBuf<uint8_t> slot;              // 9 x nel, offset of b in row of a
    DBuf<int32_t> perm;              // elements grouped by colour
    std::vector<int64_t> col_off;    // colour offsets into perm
    DBuf<double> cs;                 // vertex cos/sin (2 x nvert)
    DBuf<uint8_t> fixed;             // per dof
    std::vector<uint8_t> h_fixed;    // host copy: the radial dofs of the wall corners
    DBuf<double> E64;                // 36 x nel element matrices (rotated frame, masked)
    DBuf<float> E32;
    DBuf<double> A64, dinv64;        // 4 x nnz blocks, inverse diagonal (2 nvert)
    DBuf<float> A32, dinv32;
    double lmax64 = 0, lmax32 = 0;   // Chebyshev upper bounds of D^-1 A
    // transfer from this level (coarse) to the next finer level:
    Pattern P_pat;                   // fine node -> coarse vertex (BSR 2x2)
    DBuf<double> P64;  DBuf<float> P32;
    Pattern R_pat;                   // coarse vertex -> fine node (transpose)
    DBuf<double> R64;  DBuf<float> R32;
    // R A of the next finer P1 level (levels 1 and coarser): coarse vertex rows, fine vertex
    // columns; the V-cycle forms this level's right-hand side R (b - A x) as R b - (R A) x in
    // one launch.  Stored block q sums the (R block, A block) position pairs
    // ra_pair[2 k], ra_pair[2 k + 1] for k in [ra_ptr[q], ra_ptr[q + 1]).
    Pattern RA_pat;
    DBuf<int32_t> ra_ptr, ra_pair;
    DBuf<double> RA64; DBuf<float> RA32;
    DBuf<int64_t> fixlist;           // fixed dofs of this level
    int64_t nfix = 0;
    MGLevelBufs<double> w64;
    MGLevelBufs<float> w32;
};

// Work vectors of the thermal step, kept from one step to the next.
struct ThermalWork {
    DBuf<double> b, xf, t, r, z, p, Ap, dinv;
    DBuf<uint8_t> flag;
};
// Device copies of stk_thermal_solve's host arrays (temperature, density, fixed nodes and values).
struct ThermalIo {
    DBuf<double> t, r, vf;
    DBuf<int64_t> f;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux_stream = nullptr;   // set-up work that runs beside the context stream's
    hipStream_t dense_stream = nullptr; // second stream of the coarsest inverse (spd_inverse)
    hipEvent_t dense_ev = nullptr, dense_ev2 = nullptr;
    std::string err;

    int64_t nel = 0, nnod = 0, nvert = 0, nU = 0, nP = 0;

    // Algorithmic bytes of the kernels launched while `on` (one stk_solve): sparse and dense
    // products and Gauss-Seidel sweeps by their operands, element-wise kernels by the bytes
    // per element their call site states, or 8 bytes per element (`floor_launches`) if none.
    struct Ledger {
        bool on = false;
        double bytes = 0.0;
        int64_t launches = 0, floor_launches = 0;
        void add(double b) { if (on) { bytes += b; ++launches; } }
        void add_floor(int64_t n) { if (on) { bytes += 8.0 * (double)n; ++launches; ++floor_launches; } }
    } ledger;
    // host mesh
    std::vector<int32_t> e2n;        // 6 x nel, 0-based, element-major (e*6 + a)
    std::vector<double> el_th, el_r; // 3 x nel, element-major
    std::vector<double> node_th, node_r;
    std::vector<uint8_t> fixed;      // per dof
    std::vector<int64_t> fixed_list; // 0-based

    // colouring
    std::vector<int32_t> perm;       // elements sorted by colour
    std::vector<int64_t> col_off;    // ncolour + 1

    // patterns
    Pattern KN, GN, GTN, MN;

    // Gather lists of the element assembly: for every stored position of an operator, the element
    // blocks that sum into it, in element order (CSR over positions; codes name the element block).
    struct GatherList {
        int64_t npos = 0;
        DBuf<int32_t> ptr;
        DBuf<uint32_t> list;
    };
    GatherList gK, gG, gGT, gM, gF, gA1, gRK;
    DBuf<double> el64; DBuf<float> el32;      // per-element blocks of the last assembly
    int64_t field_cols = 6;                   // columns of eta.ip and of the tensor arrays on the device

    // device mesh data (colour-permuted, SoA: [j * nel + p])
    DBuf<int32_t> d_perm, d_e2n;
    DBuf<double> d_elth, d_elr, d_nodecs, d_noder;
    DBuf<uint8_t> d_sK, d_sG, d_sGT, d_sM, d_fixed;

    // fabric (learned preconditioner inputs) and element data
    bool have_fabric = false;
    DBuf<double> d_phi, d_delta;     // director angle, normal/shear ratio per element
    DBuf<int32_t> d_iperm;           // element -> colour position
    double mean_area = 1.0;

    // learned model
    struct MLP *mlp = nullptr;

    // fields
    bool have_fields = false;
    double g = 1.0;
    DBuf<double> d_rho, d_eta, d_C;  // eta: 6 x nel (ip-major), C: 6 comps x 6 ips x nel
    bool have_C = false;

    // operators
    int asm_flags = 0;
    bool assembled = false, loaded = false, have_mg64 = false, have_mg32 = false;
    DBuf<double> K64, G64, GT64, M64, F64, pc64;
    DBuf<float> K32, G32, GT32, M32, F32, pc32;
    // R K: the P2 -> P1 restriction times the constrained K, 2x2 blocks on the G^T pattern
    // (vertex rows, node columns); the V-cycle forms the coarse right-hand side
    // R (b - K x) as R b - (R K) x, a product with a quarter of K's rows
    DBuf<double> RK64; DBuf<float> RK32;
    DBuf<double> S64; DBuf<float> S32;        // learned Schur sum (pattern MN)
    DBuf<float> S32s;                          // its FP32 values in the sliced layout of mg[0].pat
    DBuf<double> Sdinv64; DBuf<float> Sdinv32;
    DBuf<double> Kdinv64; DBuf<float> Kdinv32; // inverse diagonal of K (per dof)
    // FP16 copy of the scaled K^ = S^-1 K S^-1 (S^2 = diag K, layout of K) for the
    // fine-level multigrid smoother, with S, S^-1 and its scaled-space work vectors
    struct HalfSmoother {
        bool on = false;
        DBuf<uint16_t> A;
        DBuf<float> s, sinv, bh, xh, xha, xu;
    } ksm;
    // Multicolour block Gauss-Seidel smoother of the fine level (FP32): K's node rows grouped
    // by colour (no two nodes of a colour share an element), each colour in slices of 64
    // rows of about equal length, blocks ordered lower-colour columns first.  A forward
    // sweep from zero reads only those (nlow per row, slow per slice); a full sweep reads
    // every block.  dinv: inverse 2x2 diagonal block per coloured row.  col holds coloured
    // row positions q: the smoother's iterate is kept in coloured row order.
    struct GSSmoother {
        bool on = false;
        int ncolour = 0;
        std::vector<int8_t> node_colour;       // colour of each node row
        std::vector<int64_t> slice_off;        // first slice of each colour (ncolour + 1)
        std::vector<int64_t> row_off;          // first coloured row of each colour (ncolour + 1)
        DBuf<int64_t> sptr;                    // slice block offsets (nslices + 1)
        DBuf<int32_t> col, src;                // per stored block: node column, position in K
        DBuf<int32_t> rowid, slow;             // per coloured row: node (-1: padding); per slice
        DBuf<uint8_t> nlow;                    // per coloured row
        DBuf<int64_t> diagpos;                 // per node: position of its diagonal block in K
        DBuf<float> val, dinv;                 // 4 per stored block; 4 per coloured row
        int64_t nstore = 0;
        // iterate and right-hand side in coloured row order (2 per coloured row), the
        // near-null vector in that order, R K with q columns, P with q rows (+ its values)
        DBuf<float> xq, bq, wq;
        Pattern rkq, pq;
        DBuf<int32_t> pq_src;
        DBuf<float> pq_val;
        bool filled = false;                   // values and dinv of the current FP32 K
        // per colour: rows, structural blocks, and blocks in lower colours (what a sweep from
        // zero reads); distinct neighbour nodes of its rows (of lower colours)
        std::vector<int64_t> rows, blocks, blocks_low, nbr, nbr_low;
    } gs;
    int fine_smoother = 0;                     // STK_SMOOTH_* of the last stk_solve
    // Incomplete Cholesky M ~ L D L^T of the pressure mass matrix (STK_PPC_IC), IC(0) on M's
    // pattern in a multicolour vertex order: vertices of one colour are not coupled, so each
    // triangular solve is one launch per colour.  Rows of L (forward) and of L^T (backward) are
    // stored by position in that order; their columns are vertex ids.
    struct PressureIC {
        bool valid = false;
        double shift = 0;                      // diagonal shift alpha of the last factorisation
        std::vector<int64_t> colour_off;       // ncolour + 1 offsets into order
        DBuf<int32_t> order;                   // vertices by colour
        DBuf<int32_t> lptr, lcol, uptr, ucol;
        DBuf<double> lval, uval, dinv;         // dinv per vertex
    } pic;
    DBuf<double> rot64; DBuf<float> rot32;     // unit near-null vector of K (discrete rigid rotation)
    double null_lambda = 0;                    // its Rayleigh quotient v^T K v
    DBuf<int64_t> d_fixlist;
    double Klmax64 = 0, Klmax32 = 0;
    double fnorm = 1.0;

    // multigrid
    bool structured = false;
    int32_t mesh_split = 0;                   // structured annulus: 2 (Grid_split1) or 4 (Grid) triangles per cell
    bool rotation_free = true;                // no tangential dof fixed: the rotation is admissible
    std::string mg_note;                      // why the multigrid is not built (when it is not)
    int32_t ncell_r = 0, ncell_t = 0;
    std::vector<MGLevel> mg;                  // mg[0] = P1 on the original mesh
    bool mg_built = false;
    DBuf<double> coarse_inv64;                // dense inverse of the coarsest operator
    DBuf<float> coarse_inv32;
    // work space kept from one call to the next (same sizes every time step)
    DBuf<double> dense_A, dense_Y;            // the padded coarsest matrix and L^-1 (spd_inverse)
    DBuf<int32_t> dense_info;
    DBuf<double> nv_y, nv_d, nv_r, nv_keep;   // near-null vector set-up
    DBuf<float> nv_r32, nv_e32;
    DBuf<double> res_u, res_p;                // residuals of the solve's refinement
    DBuf<float> res_u32, res_p32;
    ThermalWork thw;
    ThermalIo thw_io;
    // Device copies of the host arrays that the API calls upload and download, kept from one call to
    // the next by name (each name has one user; calls on a context do not overlap).
    std::map<std::string, DBuf<double>> stage;
    DBuf<double> &staged(const char *name) { return stage[name]; }
    int64_t coarse_n = 0;
    bool coarse64_valid = false;              // coarse_inv64 belongs to the current assembly

    // stress post-processor: fields at the stress rule, natural-order element data,
    // device Cartesian velocity of the last stk_updatevelocity
    int s_nip = 0;
    bool s_aniso = false, have_sfields = false;
    DBuf<double> d_seta, d_sC;               // nel x nip; 6 components x nel x nip
    DBuf<int32_t> d_e2n_nat;                 // 6 x nel
    DBuf<double> d_th_nat, d_r_nat;          // 3 x nel
    DBuf<double> d_Ucart;                    // nU
    // statistics: nodes one radius level inside each wall, spacings, wall radii
    DBuf<int32_t> d_itop, d_ibot;
    double nu_dr = 0, nu_dth = 0, nu_rin = 0, nu_rout = 0;
    std::string stats_note;                  // why the statistics data could not be prepared
    double *stats_scratch = nullptr;         // this device's fixed scratch array (stk_stats_noalloc)

    // fabric: finite strain ellipses of the last stk_fse, the DEM table, the viscous tensor of
    // the last stk_anisotropic_tensor
    DBuf<double> d_fse;                      // 6 planes (x1, x2, y1, y2, a1, a2) of nel x fse_nip points
    int fse_nip = 0;
    struct DemTable {
        bool set = false;
        int32_t sblk = 0, nblk = 0;          // entries per block of equal a1, number of blocks
        double w = 0;                        // viscosity cut-off
        DBuf<double> a1a2, a2a3;             // nblk block values; sblk sorted values
        DBuf<int32_t> perm;                  // sorted position -> entry within a block
        DBuf<double> tab;                    // 4 per entry: C11, C33, C55, C13
    } dem;
    DBuf<double> d_aniso;                    // 6 components x nel x 7

    // thermal diffusion operators (sliced P2 node pattern, one value per stored entry)
    bool have_thermal = false;
    DBuf<double> KT, MT, AT, FT;

    // solution and work
    DBuf<double> U, P, work64;
    DBuf<double> red64;                       // reduction partials
    DBuf<double> d_sc;                        // device scalars
    // Scalars the host reads (residual norms): a one-thread kernel copies them from d_sc into this
    // fine-grained (coherent) host memory and then writes a sequence number, which the host polls
    double *h_pub = nullptr;
    unsigned *h_seq = nullptr;
    unsigned pub_seq = 0;
    MGLevelBufs<double> kw64, cg64;
    MGLevelBufs<float> kw32, cg32;
    SchurBufs<double> sw64;
    SchurBufs<float> sw32;

    ~Context();
};

// ---------------------------------------------------------------- set-up (setup.cpp)
void build_patterns(Context &c);
void make_sliced(Pattern &p, hipStream_t s);
// 2x2 operators with at least this many block rows use the sliced layout.
constexpr int64_t kSellMinRows = 16384;
// Multigrid coarsening stops at the first P1 lattice with at most this many vertices; that
// level is solved with a dense inverse (2 dofs per vertex).
constexpr int64_t kDenseMaxVerts = 2300;
void build_colouring(Context &c);
void build_slot_maps(Context &c);
void build_gs_layout(Context &c);
void detect_structured(Context &c);
// Gather lists of the element assembly (after the patterns and the multigrid structure).
void build_gather_lists(Context &c);
bool fixed_set_is_free_slip(const Context &c);
// In-place inverse of the SPD matrix A (np x np column-major, np a multiple of 64) with np x np work
// space Y; *info (device) receives the first non-positive pivot (1-based) or stays 0 (dense.hip).
// Work on stream s and, for the forward stage, s2 (ordered by the events ev and ev2); the result
// is ready in stream order on s.
void spd_inverse(hipStream_t s, hipStream_t s2, hipEvent_t ev, hipEvent_t ev2, double *A, double *Y, int64_t np,
                 int *info);
// IC(0) factorisation of the full pressure mass matrix M (values on MN in CSR order) for STK_PPC_IC.
void build_pressure_ic(Context &c, const std::vector<double> &M);
// Values of caller operators (1-based CSC) scattered into the layouts of K, G, G^T, M and F.
struct CscIn {
    const int64_t *colptr, *rowval;
    const double *nzval;
};
void import_operators(const Context &c, bool K_lower, CscIn K, bool M_lower, CscIn M, CscIn G, const double *Rhs,
                      std::vector<double> &Kv, std::vector<double> &Gv, std::vector<double> &GTv,
                      std::vector<double> &Mv, std::vector<double> &Fv);
void build_mg_structure(Context &c);
// R A pattern and pair lists of the levels from 1 on (called by build_mg_structure).
void build_level_products(Context &c);

// ---------------------------------------------------------------- kernels (assembly.hip)
struct SchurSampleHost {
    int64_t nsamp = 0, per_snap = 1, first_snap = 0;
    int mode = 0, ratio_kind = 1;
    float *X = nullptr, *Tg = nullptr;
    double *S = nullptr, *scale = nullptr;
};
void launch_schur_samples(Context &c, const SchurSampleHost &h);
void launch_learned_scatter(Context &c, const float *Y, const double *scale, double *S64);
template <typename T>
void assemble_impl(Context &c, int flags, T *K, T *G, T *GT, T *M, T *F, T *A1, T *E1, T *RK = nullptr);
template <typename T>
void galerkin_impl(Context &c, size_t l, T *Ef, T *Ec, T *A);
template <typename T>
void fix_diag(Context &c, const Pattern &p, int64_t nfix, const int64_t *dofs, T *A);

}  // namespace stk
