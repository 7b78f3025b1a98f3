// Host set-up of the library (dimspers_amd/csrc/setup.cpp) driven on the CPU, for an AddressSanitizer
// + UndefinedBehaviorSanitizer build with g++ against the host-memory HIP stand-in of
// tests/host_sanitize/hip/.  It fills a context as stk_create does, runs every once-per-mesh step
// (patterns, element colouring, slot maps, structure detection, multigrid structure, Gauss-Seidel
// layout, gather lists, pressure IC) and, with "operators", the CSC import of stk_set_operators;
// then checks the invariants those steps promise.  Prints one line "ok ..." or exits non-zero.
//
// Usage: setup_driver DIR [operators]   (DIR as written by oracle/cpu_port.write_inputs; with
// "operators": Kcolptr.bin ... Rhs.bin and ops.txt "nK nM nG k_lower m_lower" as tests/c_abi uses)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "stk_internal.h"

namespace stk {
// the device-side teardown of a context lives in the HIP sources; buffers free themselves
Context::~Context() {}
}  // namespace stk

using namespace stk;

template <typename T>
static std::vector<T> load(const std::string &dir, const std::string &name, size_t n) {
    std::vector<T> v(n);
    FILE *f = std::fopen((dir + "/" + name).c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(T), n, f) != n) {
        std::fprintf(stderr, "cannot read %s\n", name.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

// column node `col` present in row node `row` of the pattern
static bool has_block(const Pattern &p, int64_t row, int32_t col) {
    for (int32_t k = p.ptr[row]; k < p.ptr[row + 1]; ++k)
        if (p.col[k] == col) return true;
    return false;
}

#define REQUIRE(cond)                                                          \
    do {                                                                       \
        if (!(cond)) {                                                         \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
            return 1;                                                          \
        }                                                                      \
    } while (0)

// DBuf keeps the allocation that release() drops and gives it back to an alloc() of the same size;
// purge() and the destructor free it (the sanitizer's leak check covers the rest).
static int check_dbuf() {
    DBuf<double> b;
    b.alloc(100);
    REQUIRE(b && b.n == 100);
    double *p0 = b.p;
    b.release();
    REQUIRE(!b && b.n == 0 && b.p == nullptr);
    b.alloc(100);
    REQUIRE(b.p == p0 && b.n == 100);
    b.release();
    b.alloc(50);
    REQUIRE(b && b.n == 50);
    b.p[49] = 1.0;
    b.release();
    b.release();
    b.alloc(50);
    REQUIRE(b && b.n == 50 && b.p[49] == 1.0);
    DBuf<double> moved(std::move(b));
    REQUIRE(!b && moved.n == 50);
    moved.purge();
    REQUIRE(!moved && moved.n == 0);
    moved.alloc(50);
    REQUIRE(moved && moved.n == 50);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    if (check_dbuf() != 0) return 1;
    const std::string dir = argv[1];
    const bool operators = argc > 2 && std::string(argv[2]) == "operators";
    long long nel, nnod, nvert, nfixed;
    double g;
    FILE *m = std::fopen((dir + "/meta.txt").c_str(), "r");
    if (!m || std::fscanf(m, "%lld %lld %lld %lld %lf", &nel, &nnod, &nvert, &nfixed, &g) != 5) return 2;
    std::fclose(m);
    const auto e2n = load<int64_t>(dir, "e2n.bin", 6 * (size_t)nel);
    const auto fixed = load<int64_t>(dir, "fixed.bin", (size_t)nfixed);
    Context c;
    try {
        c.nel = nel; c.nnod = nnod; c.nvert = nvert; c.nU = 2 * nnod; c.nP = nvert;
        c.e2n.resize(6 * nel);
        for (int64_t k = 0; k < 6 * nel; ++k) c.e2n[k] = (int32_t)(e2n[k] - 1);
        c.el_th = load<double>(dir, "el_theta.bin", 3 * (size_t)nel);
        c.el_r = load<double>(dir, "el_r.bin", 3 * (size_t)nel);
        c.node_th = load<double>(dir, "node_theta.bin", (size_t)nnod);
        c.node_r = load<double>(dir, "node_r.bin", (size_t)nnod);
        c.fixed.assign(2 * nnod, 0);
        for (int64_t d : fixed) c.fixed[d - 1] = 1;
        for (int64_t d = 0; d < 2 * nnod; ++d)
            if (c.fixed[d]) c.fixed_list.push_back(d);
        build_patterns(c);
        build_colouring(c);
        build_slot_maps(c);
        detect_structured(c);
        const bool free_slip = fixed_set_is_free_slip(c);
        if (c.structured && free_slip) build_mg_structure(c);
        if (!c.mg.empty()) build_gs_layout(c);
        build_gather_lists(c);
        // an SPD matrix on M's pattern: graph Laplacian + 2 I
        std::vector<double> Mv(c.MN.nnz());
        for (int64_t i = 0; i < c.MN.nrows; ++i) {
            const int32_t deg = c.MN.ptr[i + 1] - c.MN.ptr[i];
            for (int32_t k = c.MN.ptr[i]; k < c.MN.ptr[i + 1]; ++k)
                Mv[k] = c.MN.col[k] == i ? (double)deg + 1.0 : -1.0;
        }
        build_pressure_ic(c, Mv);
    } catch (const Error &e) {
        std::fprintf(stderr, "set-up failed: %s\n", e.what());
        return 1;
    }
    // element colouring: no two elements of a colour share a node
    REQUIRE((int64_t)c.perm.size() == nel && c.col_off.back() == nel);
    for (size_t k = 0; k + 1 < c.col_off.size(); ++k) {
        std::vector<uint8_t> seen(nnod, 0);
        for (int64_t p = c.col_off[k]; p < c.col_off[k + 1]; ++p)
            for (int a = 0; a < 6; ++a) {
                const int32_t n = c.e2n[(int64_t)c.perm[p] * 6 + a];
                REQUIRE(!seen[n]);
                seen[n] = 1;
            }
    }
    // the pattern holds every element node pair
    for (int64_t e = 0; e < nel; ++e)
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b) REQUIRE(has_block(c.KN, c.e2n[e * 6 + a], c.e2n[e * 6 + b]));
    if (!c.mg.empty()) {
        // Gauss-Seidel colours: no two coupled nodes share one; every node sits in exactly one row
        const Pattern &K = c.KN;
        for (int64_t i = 0; i < K.nrows; ++i)
            for (int32_t k = K.ptr[i]; k < K.ptr[i + 1]; ++k)
                if (K.col[k] != i) REQUIRE(c.gs.node_colour[i] != c.gs.node_colour[K.col[k]]);
        int64_t rows = 0, blocks = 0;
        for (int k = 0; k < c.gs.ncolour; ++k) { rows += c.gs.rows[k]; blocks += c.gs.blocks[k]; }
        REQUIRE(rows == nnod && blocks == K.nnz());
        // the coloured-column copies of R K and P keep every structural block
        REQUIRE(c.gs.rkq.nnz() == c.GTN.nnz() && c.gs.pq.nnz() == c.mg[0].P_pat.nnz());
        // levels: sizes halve, fixed dofs are radial
        for (size_t l = 0; l < c.mg.size(); ++l) {
            const MGLevel &L = c.mg[l];
            REQUIRE((int64_t)L.h_fixed.size() == 2 * L.nvert && L.pat.nrows == L.nvert);
            for (int64_t d = 0; d < 2 * L.nvert; ++d) REQUIRE(!L.h_fixed[d] || d % 2 == 1);
            if (l > 0) REQUIRE(L.P_pat.nrows == c.mg[l - 1].nvert && L.P_pat.ncols == L.nvert);
            if (l > 0) {
                // R A: one pair per (R block, A block) product, every pair inside both value arrays,
                // and a stored block's pairs all end in its own column
                const Pattern &Q = L.RA_pat, &R = L.R_pat, &A = c.mg[l - 1].pat;
                REQUIRE(Q.nrows == L.nvert && Q.ncols == c.mg[l - 1].nvert && Q.sell == R.sell);
                std::vector<int32_t> ptr(Q.nstore() + 1), pair;
                std::memcpy(ptr.data(), L.ra_ptr.p, ptr.size() * sizeof(int32_t));
                pair.resize(2 * (size_t)ptr.back());
                std::memcpy(pair.data(), L.ra_pair.p, pair.size() * sizeof(int32_t));
                int64_t want = 0;
                for (int64_t u = 0; u < R.nrows; ++u)
                    for (int32_t i = R.ptr[u]; i < R.ptr[u + 1]; ++i) want += A.ptr[R.col[i] + 1] - A.ptr[R.col[i]];
                REQUIRE(ptr.back() == want);
                std::vector<int32_t> acol(A.nstore());
                for (int64_t g = 0; g < A.nrows; ++g)
                    for (int32_t j = A.ptr[g]; j < A.ptr[g + 1]; ++j) acol[A.pos(g, j - A.ptr[g])] = A.col[j];
                for (int64_t u = 0; u < Q.nrows; ++u)
                    for (int32_t k = Q.ptr[u]; k < Q.ptr[u + 1]; ++k) {
                        const int64_t q = Q.pos(u, k - Q.ptr[u]);
                        REQUIRE(ptr[q] <= ptr[q + 1]);
                        for (int32_t t = ptr[q]; t < ptr[q + 1]; ++t) {
                            REQUIRE(pair[2 * t] >= 0 && pair[2 * t] < R.nstore());
                            REQUIRE(pair[2 * t + 1] >= 0 && pair[2 * t + 1] < A.nstore());
                            REQUIRE(acol[pair[2 * t + 1]] == Q.col[k]);
                        }
                    }
            }
        }
    }
    // gather lists: every element block lands in exactly one stored position
    REQUIRE(c.gK.npos == c.KN.nstore() && c.gM.npos == c.MN.nnz());
    {
        std::vector<int32_t> ptr(c.gK.npos + 1);
        std::memcpy(ptr.data(), c.gK.ptr.p, ptr.size() * sizeof(int32_t));
        REQUIRE(ptr.back() == 36 * nel);
        for (int64_t q = 0; q < c.gK.npos; ++q) REQUIRE(ptr[q] <= ptr[q + 1]);
    }
    REQUIRE(c.pic.valid);
    int64_t kvals = 0;
    if (operators) {
        long long nK, nM, nG;
        int kl, ml;
        FILE *o = std::fopen((dir + "/ops.txt").c_str(), "r");
        if (!o || std::fscanf(o, "%lld %lld %lld %d %d", &nK, &nM, &nG, &kl, &ml) != 5) return 2;
        std::fclose(o);
        const auto Kp = load<int64_t>(dir, "Kcolptr.bin", 2 * nnod + 1), Kr = load<int64_t>(dir, "Krowval.bin", nK);
        const auto Kz = load<double>(dir, "Knzval.bin", nK);
        const auto Mp = load<int64_t>(dir, "Mcolptr.bin", nvert + 1), Mr = load<int64_t>(dir, "Mrowval.bin", nM);
        const auto Mz = load<double>(dir, "Mnzval.bin", nM);
        const auto Gp = load<int64_t>(dir, "Gcolptr.bin", nvert + 1), Gr = load<int64_t>(dir, "Growval.bin", nG);
        const auto Gz = load<double>(dir, "Gnzval.bin", nG);
        const auto Rhs = load<double>(dir, "Rhs.bin", 2 * nnod);
        std::vector<double> Kv, Gv, GTv, Mv, Fv;
        try {
            import_operators(c, kl != 0, CscIn{Kp.data(), Kr.data(), Kz.data()}, ml != 0,
                             CscIn{Mp.data(), Mr.data(), Mz.data()}, CscIn{Gp.data(), Gr.data(), Gz.data()},
                             Rhs.data(), Kv, Gv, GTv, Mv, Fv);
        } catch (const Error &e) {
            std::fprintf(stderr, "import failed: %s\n", e.what());
            return 1;
        }
        // G and G^T carry the same values; a refused entry outside the pattern
        double sg = 0, sgt = 0;
        for (double x : Gv) sg += x;
        for (double x : GTv) sgt += x;
        REQUIRE(std::abs(sg - sgt) <= 1e-9 * (std::abs(sg) + 1.0));
        bool refused = false;
        try {
            std::vector<double> a, b, cc, d, e;
            std::vector<int64_t> far_r(Kr);
            far_r[0] = 2 * nnod;   // row 2 nnod (1-based) in column 1: not a row of node 0's pattern
            import_operators(c, false, CscIn{Kp.data(), far_r.data(), Kz.data()}, ml != 0,
                             CscIn{Mp.data(), Mr.data(), Mz.data()}, CscIn{Gp.data(), Gr.data(), Gz.data()},
                             Rhs.data(), a, b, cc, d, e);
        } catch (const Error &) {
            refused = true;
        }
        REQUIRE(refused);
        kvals = (int64_t)Kv.size();
    }
    std::printf("ok split=%d levels=%zu colours=%zu gs_colours=%d ic_colours=%zu kvals=%lld\n", c.mesh_split,
                c.mg.size(), c.col_off.size() - 1, c.gs.ncolour, c.pic.colour_off.size() - 1, (long long)kvals);
    return 0;
}
