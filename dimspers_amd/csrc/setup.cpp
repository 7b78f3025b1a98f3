// Host-side, once-per-mesh set-up: node-block sparsity patterns, element colouring,
// per-element scatter slots and the multigrid hierarchy structure.
//
// The reference builds triplet index lists once per mesh (sparsitystokes,
// src/Algebra/Sparsity.jl:10-103) and lets sparse(I, J, V) sort and sum them at
// every assembly.  Here the patterns are built once as node-block CSR, and every
// element records, for each of its node pairs, the offset of the column node in
// the row node's pattern row.  Elements are coloured so that no two elements of
// one colour share a node (src/Grid/coloring.jl:5-64 colours greedily by first
// fit; the balanced variant below keeps the colours of equal size, which keeps
// every launch large).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "stk_internal.h"

namespace stk {

namespace {

// Rows: for every row node, the sorted unique column nodes reachable through elements.
void pattern_from_elements(Pattern &pat, int64_t nrows, int64_t ncols, int64_t nel,
                           const std::vector<int32_t> &e2r, int nr_per_el,
                           const std::vector<int32_t> &e2c, int nc_per_el) {
    // row -> elements incidence
    std::vector<int64_t> cnt(nrows + 1, 0);
    for (int64_t e = 0; e < nel; ++e)
        for (int a = 0; a < nr_per_el; ++a) cnt[e2r[e * nr_per_el + a] + 1]++;
    for (int64_t i = 0; i < nrows; ++i) cnt[i + 1] += cnt[i];
    std::vector<int32_t> inc(cnt[nrows]);
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < nel; ++e)
        for (int a = 0; a < nr_per_el; ++a) inc[pos[e2r[e * nr_per_el + a]]++] = (int32_t)e;

    std::vector<int32_t> deg(nrows);
    std::vector<std::vector<int32_t>> rows(nrows);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nrows; ++i) {
        std::vector<int32_t> &r = rows[i];
        r.reserve(64);
        for (int64_t k = cnt[i]; k < cnt[i + 1]; ++k) {
            int64_t e = inc[k];
            for (int b = 0; b < nc_per_el; ++b) r.push_back(e2c[e * nc_per_el + b]);
        }
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        deg[i] = (int32_t)r.size();
    }
    pat.nrows = nrows;
    pat.ncols = ncols;
    pat.ptr.assign(nrows + 1, 0);
    for (int64_t i = 0; i < nrows; ++i) pat.ptr[i + 1] = pat.ptr[i] + deg[i];
    pat.col.resize(pat.ptr[nrows]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nrows; ++i)
        std::copy(rows[i].begin(), rows[i].end(), pat.col.begin() + pat.ptr[i]);
}

inline int find_offset(const Pattern &p, int32_t row, int32_t col) {
    const int32_t *b = p.col.data() + p.ptr[row];
    const int32_t *e = p.col.data() + p.ptr[row + 1];
    const int32_t *it = std::lower_bound(b, e, col);
    if (it == e || *it != col) throw Error(STK_E_SHAPE, "pattern lookup failed");
    int off = (int)(it - b);
    if (off > 255) throw Error(STK_E_SHAPE, "pattern row longer than 256 blocks");
    return off;
}

// Balanced greedy colouring: every element takes, among the colours none of its
// nodes carries yet, the one with the fewest elements so far.
void colour_elements(int64_t nel, int nn, const std::vector<int32_t> &e2x, int64_t nnodes,
                     std::vector<int32_t> &perm, std::vector<int64_t> &off) {
    std::vector<uint64_t> used(nnodes, 0);
    std::vector<int32_t> colour(nel);
    std::vector<int64_t> count;
    for (int64_t e = 0; e < nel; ++e) {
        uint64_t forbid = 0;
        for (int a = 0; a < nn; ++a) forbid |= used[e2x[e * nn + a]];
        int best = -1;
        for (int k = 0; k < (int)count.size(); ++k)
            if (!(forbid >> k & 1ull) && (best < 0 || count[k] < count[best])) best = k;
        if (best < 0) {
            best = (int)count.size();
            if (best >= 64) throw Error(STK_E_SHAPE, "colouring needs more than 64 colours");
            count.push_back(0);
        }
        colour[e] = best;
        count[best]++;
        for (int a = 0; a < nn; ++a) used[e2x[e * nn + a]] |= 1ull << best;
    }
    int nc = (int)count.size();
    off.assign(nc + 1, 0);
    for (int k = 0; k < nc; ++k) off[k + 1] = off[k] + count[k];
    perm.resize(nel);
    std::vector<int64_t> p(off.begin(), off.end() - 1);
    for (int64_t e = 0; e < nel; ++e) perm[p[colour[e]]++] = (int32_t)e;
}

}  // namespace

void build_patterns(Context &c) {
    std::vector<int32_t> e2v(c.nel * 3);
    for (int64_t e = 0; e < c.nel; ++e)
        for (int a = 0; a < 3; ++a) e2v[e * 3 + a] = c.e2n[e * 6 + a];
    pattern_from_elements(c.KN, c.nnod, c.nnod, c.nel, c.e2n, 6, c.e2n, 6);
    pattern_from_elements(c.GN, c.nnod, c.nvert, c.nel, c.e2n, 6, e2v, 3);
    pattern_from_elements(c.GTN, c.nvert, c.nnod, c.nel, e2v, 3, c.e2n, 6);
    pattern_from_elements(c.MN, c.nvert, c.nvert, c.nel, e2v, 3, e2v, 3);
    for (Pattern *p : {&c.KN, &c.GN, &c.GTN, &c.MN}) {
        p->dptr.upload(p->ptr, c.stream);
        p->dcol.upload(p->col, c.stream);
    }
    make_sliced(c.KN, c.stream);
    make_sliced(c.GN, c.stream);
    make_sliced(c.GTN, c.stream);
}

void make_sliced(Pattern &p, hipStream_t s) {
    const int64_t C = Pattern::kSlice;
    const int64_t ns = (p.nrows + C - 1) / C;
    p.sptr.assign(ns + 1, 0);
    for (int64_t sl = 0; sl < ns; ++sl) {
        int64_t len = 0;
        for (int64_t r = sl * C; r < std::min<int64_t>(p.nrows, (sl + 1) * C); ++r)
            len = std::max<int64_t>(len, p.ptr[r + 1] - p.ptr[r]);
        p.sptr[sl + 1] = p.sptr[sl] + len * C;
    }
    p.scol.assign(p.sptr[ns], 0);
    for (int64_t sl = 0; sl < ns; ++sl) {
        const int64_t len = (p.sptr[sl + 1] - p.sptr[sl]) / C;
        for (int64_t l = 0; l < C; ++l) {
            const int64_t r = sl * C + l;
            for (int64_t k = 0; k < len; ++k) {
                const int64_t q = p.sptr[sl] + k * C + l;
                // padding (zero block): repeat the row's last column, which is a valid
                // index of x for any pattern shape (rectangular transfers included)
                const int64_t len_r = r < p.nrows ? p.ptr[r + 1] - p.ptr[r] : 0;
                if (k < len_r) p.scol[q] = p.col[p.ptr[r] + k];
                else p.scol[q] = len_r > 0 ? p.col[p.ptr[r + 1] - 1] : 0;
            }
        }
    }
    for (int32_t cidx : p.scol)
        if (cidx < 0 || cidx >= p.ncols) throw Error(STK_E_STATE, "sliced pattern column out of range");
    p.sell = true;
    p.dsptr.upload(p.sptr, s);
    p.dscol.upload(p.scol, s);
}

void build_colouring(Context &c) { colour_elements(c.nel, 6, c.e2n, c.nnod, c.perm, c.col_off); }

void build_gs_layout(Context &c) {
    const Pattern &K = c.KN;
    const int64_t n = K.nrows, C = Pattern::kSlice;
    Context::GSSmoother &g = c.gs;
    // greedy node colouring in natural order (P2 nodes: 7 colours on the annulus meshes)
    std::vector<int8_t> colour(n, -1);
    int nc = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint64_t used = 0;
        for (int32_t k = K.ptr[i]; k < K.ptr[i + 1]; ++k)
            if (colour[K.col[k]] >= 0) used |= 1ull << colour[K.col[k]];
        int ci = 0;
        while (used >> ci & 1ull) ++ci;
        if (ci >= 64) throw Error(STK_E_SHAPE, "node colouring needs more than 64 colours");
        colour[i] = (int8_t)ci;
        nc = std::max(nc, ci + 1);
    }
    // rows of each colour, longest first (homogeneous slices), then by node
    std::vector<std::vector<int32_t>> rows(nc);
    for (int64_t i = 0; i < n; ++i) rows[colour[i]].push_back((int32_t)i);
    auto len = [&](int32_t r) { return K.ptr[r + 1] - K.ptr[r]; };
    for (auto &v : rows)
        std::stable_sort(v.begin(), v.end(), [&](int32_t a, int32_t b) { return len(a) > len(b); });
    g.ncolour = nc;
    g.node_colour = colour;
    g.slice_off.assign(nc + 1, 0);
    g.row_off.assign(nc + 1, 0);
    // one lane per row (two or four lanes per row, with shorter per-lane load chains, were
    // measured no faster at 256 x 2048)
    std::vector<int32_t> rowid;
    for (int k = 0; k < nc; ++k) {
        const int64_t S = C;
        const int64_t ns = ((int64_t)rows[k].size() + S - 1) / S;
        g.slice_off[k + 1] = g.slice_off[k] + ns;
        g.row_off[k + 1] = g.row_off[k] + ns * S;
        for (int64_t q = 0; q < ns * S; ++q) rowid.push_back(q < (int64_t)rows[k].size() ? rows[k][q] : -1);
    }
    const int64_t nsl = g.slice_off[nc];
    std::vector<int64_t> sptr(nsl + 1, 0);
    std::vector<int32_t> slow(nsl, 0);
    std::vector<uint8_t> nlow(rowid.size(), 0);
    // per row: block order (indices into the CSR row), lower-colour columns first
    std::vector<int64_t> rlow(n, 0);
    std::vector<int32_t> order(K.col.size());
    for (int64_t i = 0; i < n; ++i) {
        int64_t q = K.ptr[i];
        int nl = 0;
        for (int32_t k = K.ptr[i]; k < K.ptr[i + 1]; ++k)
            if (colour[K.col[k]] < colour[i]) { order[q++] = k - K.ptr[i]; ++nl; }
        for (int32_t k = K.ptr[i]; k < K.ptr[i + 1]; ++k)
            if (colour[K.col[k]] >= colour[i]) order[q++] = k - K.ptr[i];
        if (nl > 255) throw Error(STK_E_SHAPE, "row with more than 255 lower-colour blocks");
        rlow[i] = nl;
    }
    g.rows.assign(nc, 0);
    g.blocks.assign(nc, 0);
    g.blocks_low.assign(nc, 0);
    for (int64_t i = 0; i < n; ++i) {
        g.rows[colour[i]] += 1;
        g.blocks[colour[i]] += len((int32_t)i);
        g.blocks_low[colour[i]] += rlow[i];
    }
    // the distinct neighbour nodes a colour's launch gathers x from (every one at least once per
    // launch: x changes between launches), all of them and those of lower colours (from zero)
    {
        std::vector<uint64_t> adj(n, 0);
        for (int64_t i = 0; i < n; ++i)
            for (int32_t k = K.ptr[i]; k < K.ptr[i + 1]; ++k)
                if (K.col[k] != i) adj[K.col[k]] |= 1ull << colour[i];
        g.nbr.assign(nc, 0);
        g.nbr_low.assign(nc, 0);
        for (int64_t j = 0; j < n; ++j)
            for (int k = 0; k < nc; ++k)
                if (adj[j] >> k & 1ull) {
                    g.nbr[k] += 1;
                    if (colour[j] < k) g.nbr_low[k] += 1;
                }
    }
    // slice s of colour k holds rows row_off[k] + (s - slice_off[k]) * 64 + j, j < 64;
    // block b of its row j at sptr[s] + b * 64 + j
    for (int k = 0; k < nc; ++k) {
        const int64_t S = C;
        for (int64_t s = g.slice_off[k]; s < g.slice_off[k + 1]; ++s) {
            const int64_t r0 = g.row_off[k] + (s - g.slice_off[k]) * S;
            int64_t Lb = 0, lo = 0;
            for (int64_t j = 0; j < S; ++j) {
                const int32_t r = rowid[r0 + j];
                if (r < 0) continue;
                Lb = std::max<int64_t>(Lb, len(r));
                lo = std::max<int64_t>(lo, rlow[r]);
                nlow[r0 + j] = (uint8_t)rlow[r];
            }
            sptr[s + 1] = sptr[s] + Lb * S;
            slow[s] = (int32_t)lo;
        }
    }
    g.nstore = sptr[nsl];
    std::vector<int32_t> col(g.nstore), src(g.nstore);
    for (int k = 0; k < nc; ++k) {
        const int64_t S = C;
        for (int64_t s = g.slice_off[k]; s < g.slice_off[k + 1]; ++s) {
            const int64_t r0 = g.row_off[k] + (s - g.slice_off[k]) * S;
            const int64_t Lb = (sptr[s + 1] - sptr[s]) / S;
            for (int64_t j = 0; j < S; ++j) {
                const int32_t r = rowid[r0 + j];
                for (int64_t b = 0; b < Lb; ++b) {
                    const int64_t q = sptr[s] + b * S + j;
                    if (r >= 0 && b < len(r)) {
                        const int32_t jj = order[K.ptr[r] + b];
                        col[q] = K.col[K.ptr[r] + jj];
                        src[q] = (int32_t)K.pos(r, jj);
                    } else {   // padding: zero block on a valid column
                        col[q] = r >= 0 ? K.col[K.ptr[r + 1] - 1] : 0;
                        src[q] = -1;
                    }
                }
            }
        }
    }
    std::vector<int64_t> diagpos(n);
    for (int64_t i = 0; i < n; ++i) diagpos[i] = K.pos(i, find_offset(K, (int32_t)i, (int32_t)i));
    for (int32_t x : col)
        if (x < 0 || x >= K.ncols) throw Error(STK_E_STATE, "coloured pattern column out of range");
    // the smoother's iterate lives in coloured row order (q): columns of the coloured K, of
    // R K and the rows of the prolongation are renumbered to q
    const int64_t nq = g.row_off[nc];
    std::vector<int32_t> qpos(n, -1);
    for (int64_t q = 0; q < nq; ++q)
        if (rowid[q] >= 0) qpos[rowid[q]] = (int32_t)q;
    for (int64_t i = 0; i < n; ++i)
        if (qpos[i] < 0) throw Error(STK_E_STATE, "node missing from the colouring");
    for (int32_t &x : col) x = qpos[x];
    {
        const Pattern &GT = c.GTN;
        if (!GT.sell) throw Error(STK_E_STATE, "G^T pattern is not sliced");
        Pattern &Q = g.rkq;
        Q.nrows = GT.nrows; Q.ncols = nq; Q.sell = true;
        Q.sptr = GT.sptr;
        Q.scol.resize(GT.scol.size());
        for (size_t k = 0; k < GT.scol.size(); ++k) Q.scol[k] = qpos[GT.scol[k]];
        // host CSR copy (the structural blocks: the byte counts of its products read it)
        Q.ptr = GT.ptr;
        Q.col.resize(GT.col.size());
        for (size_t k = 0; k < GT.col.size(); ++k) Q.col[k] = qpos[GT.col[k]];
        Q.dsptr.upload(Q.sptr, c.stream);
        Q.dscol.upload(Q.scol, c.stream);
    }
    {
        // prolongation rows in q order: row q is P's row of node rowid[q] (empty for padding)
        const Pattern &P = c.mg[0].P_pat;
        Pattern &Q = g.pq;
        Q.nrows = nq; Q.ncols = P.ncols; Q.sell = true;
        const int64_t ns = nq / C;
        Q.sptr.assign(ns + 1, 0);
        auto plen = [&](int32_t r) { return r < 0 ? 0 : P.ptr[r + 1] - P.ptr[r]; };
        for (int64_t s = 0; s < ns; ++s) {
            int64_t Lb = 0;
            for (int64_t j = 0; j < C; ++j) Lb = std::max<int64_t>(Lb, plen(rowid[s * C + j]));
            Q.sptr[s + 1] = Q.sptr[s] + Lb * C;
        }
        Q.scol.assign(Q.sptr[ns], 0);
        std::vector<int32_t> psrc(Q.sptr[ns], -1);
        for (int64_t s = 0; s < ns; ++s) {
            const int64_t Lb = (Q.sptr[s + 1] - Q.sptr[s]) / C;
            for (int64_t j = 0; j < C; ++j) {
                const int32_t r = rowid[s * C + j];
                const int64_t lr = plen(r);
                for (int64_t b = 0; b < Lb; ++b) {
                    const int64_t qq = Q.sptr[s] + b * C + j;
                    if (b < lr) {
                        Q.scol[qq] = P.col[P.ptr[r] + b];
                        psrc[qq] = (int32_t)P.pos(r, b);
                    } else {
                        Q.scol[qq] = lr > 0 ? P.col[P.ptr[r + 1] - 1] : 0;
                    }
                }
            }
        }
        for (int32_t x : Q.scol)
            if (x < 0 || x >= Q.ncols) throw Error(STK_E_STATE, "coloured prolongation column out of range");
        Q.ptr.assign(nq + 1, 0);
        for (int64_t q = 0; q < nq; ++q) Q.ptr[q + 1] = Q.ptr[q] + plen(rowid[q]);
        Q.col.resize(Q.ptr[nq]);
        for (int64_t q = 0; q < nq; ++q)
            if (rowid[q] >= 0)
                std::copy(P.col.begin() + P.ptr[rowid[q]], P.col.begin() + P.ptr[rowid[q] + 1], Q.col.begin() + Q.ptr[q]);
        Q.dsptr.upload(Q.sptr, c.stream);
        Q.dscol.upload(Q.scol, c.stream);
        g.pq_src.upload(psrc, c.stream);
    }
    g.sptr.upload(sptr, c.stream);
    g.col.upload(col, c.stream);
    g.src.upload(src, c.stream);
    g.rowid.upload(rowid, c.stream);
    g.slow.upload(slow, c.stream);
    g.nlow.upload(nlow, c.stream);
    g.diagpos.upload(diagpos, c.stream);
    STK_HIP(hipStreamSynchronize(c.stream));
}

void build_slot_maps(Context &c) {
    const int64_t nel = c.nel;
    std::vector<uint8_t> sK(36 * nel), sG(18 * nel), sGT(18 * nel), sM(9 * nel);
    std::vector<int32_t> e2n_p(6 * nel);
    std::vector<double> th_p(3 * nel), r_p(3 * nel);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < nel; ++p) {
        const int64_t e = c.perm[p];
        const int32_t *n = &c.e2n[e * 6];
        for (int a = 0; a < 6; ++a) {
            e2n_p[a * nel + p] = n[a];
            for (int b = 0; b < 6; ++b) sK[(a * 6 + b) * nel + p] = (uint8_t)find_offset(c.KN, n[a], n[b]);
            for (int i = 0; i < 3; ++i) {
                sG[(a * 3 + i) * nel + p] = (uint8_t)find_offset(c.GN, n[a], n[i]);
                sGT[(i * 6 + a) * nel + p] = (uint8_t)find_offset(c.GTN, n[i], n[a]);
            }
        }
        for (int i = 0; i < 3; ++i) {
            th_p[i * nel + p] = c.el_th[e * 3 + i];
            r_p[i * nel + p] = c.el_r[e * 3 + i];
            for (int j = 0; j < 3; ++j) sM[(i * 3 + j) * nel + p] = (uint8_t)find_offset(c.MN, n[i], n[j]);
        }
    }
    std::vector<int32_t> perm32(c.perm.begin(), c.perm.end());
    c.d_perm.upload(perm32, c.stream);
    std::vector<int32_t> iperm(nel);
    double area = 0;
    for (int64_t p = 0; p < nel; ++p) iperm[c.perm[p]] = (int32_t)p;
    for (int64_t e = 0; e < nel; ++e) {
        const double *t = &c.el_th[3 * e], *r = &c.el_r[3 * e];
        area += ((t[1] - t[0]) * (r[2] - r[0]) - (t[2] - t[0]) * (r[1] - r[0])) * (r[0] + r[1] + r[2]) / 6.0;
    }
    c.mean_area = area / (double)nel;
    c.d_iperm.upload(iperm, c.stream);
    c.d_e2n.upload(e2n_p, c.stream);
    c.d_elth.upload(th_p, c.stream);
    c.d_elr.upload(r_p, c.stream);
    c.d_sK.upload(sK, c.stream);
    c.d_sG.upload(sG, c.stream);
    c.d_sGT.upload(sGT, c.stream);
    c.d_sM.upload(sM, c.stream);
    std::vector<double> cs(2 * c.nnod);
    for (int64_t i = 0; i < c.nnod; ++i) {
        cs[2 * i] = std::cos(c.node_th[i]);
        cs[2 * i + 1] = std::sin(c.node_th[i]);
    }
    c.d_nodecs.upload(cs, c.stream);
    c.d_noder.upload(c.node_r, c.stream);
    c.d_fixed.upload(c.fixed, c.stream);
    STK_HIP(hipStreamSynchronize(c.stream));
}

namespace {

// CSR of contributions over npos stored positions: for each element e (in order) and each of its
// `per_el` local blocks j, pos(e, j) names the position and code(e, j) the source block.  Filling
// element by element keeps every position's list in element order (a fixed summation order).
template <typename PosF, typename CodeF>
void build_list(Context::GatherList &g, int64_t npos, int64_t nel, int per_el, PosF pos, CodeF code,
                hipStream_t s) {
    std::vector<int32_t> ptr(npos + 1, 0);
    std::vector<int64_t> where((size_t)nel * per_el);
    for (int64_t e = 0; e < nel; ++e)
        for (int j = 0; j < per_el; ++j) {
            const int64_t q = pos(e, j);
            if (q < 0 || q >= npos) throw Error(STK_E_STATE, "gather list position out of range");
            where[(size_t)e * per_el + j] = q;
            ptr[q + 1]++;
        }
    for (int64_t q = 0; q < npos; ++q) ptr[q + 1] += ptr[q];
    std::vector<uint32_t> list(ptr[npos]);
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t e = 0; e < nel; ++e)
        for (int j = 0; j < per_el; ++j) list[fill[where[(size_t)e * per_el + j]]++] = code(e, j);
    g.npos = npos;
    g.ptr.upload(ptr, s);
    g.list.upload(list, s);
}

}  // namespace

void build_gather_lists(Context &c) {
    const int64_t nel = c.nel;
    // codes index element blocks [block][nel]; K's carry a transposition bit
    if ((uint64_t)2 * 21 * (uint64_t)nel >= (uint64_t)UINT32_MAX) throw Error(STK_E_SHAPE, "mesh too large for the gather codes");
    const int32_t *n = c.e2n.data();   // element-major: n[e * 6 + a]
    auto lb = [](int a, int b) { return a * (a + 1) / 2 + b; };
    // K: node pair (a, b) -> block row n_a, column n_b; the lower block (max, min) of K_e, transposed
    // when a < b
    build_list(c.gK, c.KN.nstore(), nel, 36,
               [&](int64_t e, int j) {
                   const int a = j / 6, b = j % 6;
                   const int32_t na = n[e * 6 + a], nb = n[e * 6 + b];
                   return c.KN.pos(na, find_offset(c.KN, na, nb));
               },
               [&](int64_t e, int j) {
                   const int a = j / 6, b = j % 6;
                   const int k = a >= b ? lb(a, b) : lb(b, a);
                   return (uint32_t)(((uint64_t)k * nel + e) * 2 + (a < b ? 1 : 0));
               },
               c.stream);
    // G (node n_a, vertex n_i) and G^T (vertex n_i, node n_a): block a * 3 + i of G_e
    auto gcode = [&](int64_t e, int j) { return (uint32_t)((uint64_t)j * nel + e); };
    build_list(c.gG, c.GN.nstore(), nel, 18,
               [&](int64_t e, int j) {
                   const int32_t na = n[e * 6 + j / 3], ni = n[e * 6 + j % 3];
                   return c.GN.pos(na, find_offset(c.GN, na, ni));
               },
               gcode, c.stream);
    build_list(c.gGT, c.GTN.nstore(), nel, 18,
               [&](int64_t e, int j) {
                   const int32_t na = n[e * 6 + j / 3], ni = n[e * 6 + j % 3];
                   return c.GTN.pos(ni, find_offset(c.GTN, ni, na));
               },
               gcode, c.stream);
    // M (vertex i, vertex j): the upper entry (min, max) of M_e
    static const int mpair[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    build_list(c.gM, c.MN.nnz(), nel, 9,
               [&](int64_t e, int j) {
                   const int32_t ni = n[e * 6 + j / 3], nj = n[e * 6 + j % 3];
                   return (int64_t)c.MN.ptr[ni] + find_offset(c.MN, ni, nj);
               },
               [&](int64_t e, int j) { return (uint32_t)((uint64_t)mpair[j / 3][j % 3] * nel + e); }, c.stream);
    // F (node n_a): block a of f_e
    build_list(c.gF, c.nnod, nel, 6, [&](int64_t e, int j) { return (int64_t)n[e * 6 + j]; }, gcode, c.stream);
    if (!c.mg.empty()) {
        // the P1 operator (vertex u, vertex v) on the level-0 pattern and R K (vertex u, node f) on G^T's
        const Pattern &P0 = c.mg[0].pat;
        build_list(c.gA1, P0.nstore(), nel, 9,
                   [&](int64_t e, int j) {
                       const int32_t nu = n[e * 6 + j / 3], nv = n[e * 6 + j % 3];
                       return P0.pos(nu, find_offset(P0, nu, nv));
                   },
                   gcode, c.stream);
        build_list(c.gRK, c.GTN.nstore(), nel, 18,
                   [&](int64_t e, int j) {
                       const int32_t nu = n[e * 6 + j / 6], nf = n[e * 6 + j % 6];
                       return c.GTN.pos(nu, find_offset(c.GTN, nu, nf));
                   },
                   gcode, c.stream);
    }
    STK_HIP(hipStreamSynchronize(c.stream));
}

// The multigrid needs one of the structured annuli of src/Grid/mesher.jl, both with corner
// vertex v(i, j) = i + nr*j (radial fastest) and cell q = i + Nr*j:
//  * Grid_split1 (:152-233): elements 2q and 2q+1 with vertices [v(i,j), v(i,j+1), v(i+1,j)]
//    and [v(i,j+1), v(i+1,j+1), v(i+1,j)];
//  * Grid (:59-150, init_grid(N; split=2), the mesh MAIN.jl:33 builds): centre vertex nr*nt + q,
//    elements 4q + t = [K_t, K_t+1, centre] with the corners K = v(i,j), v(i,j+1), v(i+1,j+1),
//    v(i+1,j) (:88-115).
void detect_structured(Context &c) {
    c.structured = false;
    c.mesh_split = 0;
    // number of radial vertices: run of equal angles at the start of the vertex list
    int64_t nr = 1;
    while (nr < c.nvert && c.node_th[nr] == c.node_th[0]) ++nr;
    if (nr < 2) return;
    const int64_t ncr = nr - 1;
    auto split2 = [&]() {
        if (c.nvert % nr != 0) return false;
        const int64_t nt = c.nvert / nr;
        if (c.nel != 2 * ncr * nt) return false;
        for (int64_t q = 0; q < ncr * nt; ++q) {
            const int64_t i = q % ncr, j = q / ncr, jp = (j + 1) % nt;
            const int32_t v00 = (int32_t)(i + nr * j), v01 = (int32_t)(i + nr * jp);
            const int32_t v10 = v00 + 1, v11 = v01 + 1;
            const int32_t *a = &c.e2n[(2 * q) * 6], *b = &c.e2n[(2 * q + 1) * 6];
            if (a[0] != v00 || a[1] != v01 || a[2] != v10) return false;
            if (b[0] != v01 || b[1] != v11 || b[2] != v10) return false;
        }
        c.ncell_t = (int32_t)nt;
        return true;
    };
    auto split4 = [&]() {
        if (c.nvert % (nr + ncr) != 0) return false;
        const int64_t nt = c.nvert / (nr + ncr);
        if (c.nel != 4 * ncr * nt) return false;
        for (int64_t q = 0; q < ncr * nt; ++q) {
            const int64_t i = q % ncr, j = q / ncr, jp = (j + 1) % nt;
            const int32_t k[4] = {(int32_t)(i + nr * j), (int32_t)(i + nr * jp), (int32_t)(i + nr * jp + 1),
                                  (int32_t)(i + nr * j + 1)};
            const int32_t ctr = (int32_t)(nr * nt + q);
            for (int t = 0; t < 4; ++t) {
                const int32_t *a = &c.e2n[(4 * q + t) * 6];
                if (a[0] != k[t] || a[1] != k[(t + 1) % 4] || a[2] != ctr) return false;
            }
        }
        c.ncell_t = (int32_t)nt;
        return true;
    };
    if (split2()) c.mesh_split = 2;
    else if (split4()) c.mesh_split = 4;
    else return;
    c.ncell_r = (int32_t)ncr;
    c.structured = true;
}

// True when the fixed dofs are exactly the radial dofs of the nodes on the smallest and the
// largest node radius (point_ids, src/Grid/mesher.jl:419-441, compares radii for equality).
bool fixed_set_is_free_slip(const Context &c) {
    if (c.nnod == 0) return false;
    const auto mm = std::minmax_element(c.node_r.begin(), c.node_r.end());
    const double rmin = *mm.first, rmax = *mm.second;
    for (int64_t n = 0; n < c.nnod; ++n) {
        const bool wall = c.node_r[n] == rmin || c.node_r[n] == rmax;
        if (c.fixed[2 * n] != 0 || (c.fixed[2 * n + 1] != 0) != wall) return false;
    }
    return true;
}

namespace {

// Index k of block column `col` in row `row` of p (columns sorted), or -1.
int64_t find_block(const Pattern &p, int64_t row, int64_t col) {
    const int32_t *b = p.col.data() + p.ptr[row], *e = p.col.data() + p.ptr[row + 1];
    const int32_t *it = std::lower_bound(b, e, (int32_t)col);
    return (it == e || *it != (int32_t)col) ? -1 : (int64_t)(it - b);
}

// Walks the entries of a 1-based CSC matrix after checking its structure.
template <typename F>
void for_each_csc(const CscIn &a, int64_t nrows, int64_t ncols, const char *name, F f) {
    STK_CHECK(a.colptr && a.rowval && a.nzval, STK_E_ARG, std::string("NULL array of ") + name);
    STK_CHECK(a.colptr[0] == 1, STK_E_SHAPE, std::string(name) + ": colptr[0] must be 1");
    for (int64_t j = 0; j < ncols; ++j) {
        STK_CHECK(a.colptr[j + 1] >= a.colptr[j], STK_E_SHAPE, std::string(name) + ": colptr decreases");
        for (int64_t k = a.colptr[j] - 1; k < a.colptr[j + 1] - 1; ++k) {
            const int64_t i = a.rowval[k] - 1;
            STK_CHECK(i >= 0 && i < nrows, STK_E_SHAPE, std::string(name) + ": row index out of range");
            f(i, j, a.nzval[k]);
        }
    }
}

}  // namespace

void import_operators(const Context &c, bool K_lower, CscIn K, bool M_lower, CscIn M, CscIn G, const double *Rhs,
                      std::vector<double> &Kv, std::vector<double> &Gv, std::vector<double> &GTv,
                      std::vector<double> &Mv, std::vector<double> &Fv) {
    auto outside = [](const char *name, int64_t i, int64_t j) {
        return Error(STK_E_SHAPE, std::string(name) + " entry (" + std::to_string(i + 1) + ", " +
                                      std::to_string(j + 1) + ") lies outside the mesh's pattern");
    };
    Kv.assign(4 * (size_t)c.KN.nstore(), 0.0);
    auto put_k = [&](int64_t i, int64_t j, double v) {
        const int64_t k = find_block(c.KN, i / 2, j / 2);
        if (k < 0) throw outside("K", i, j);
        if (c.fixed[i] || c.fixed[j]) return;          // K[ifree, ifree] only; identity set later
        Kv[4 * c.KN.pos(i / 2, k) + (i % 2) * 2 + (j % 2)] += v;
    };
    for_each_csc(K, c.nU, c.nU, "K", [&](int64_t i, int64_t j, double v) {
        if (K_lower) {
            STK_CHECK(i >= j, STK_E_SHAPE, "K: entry above the diagonal with K_lower");
            if (i != j) put_k(j, i, v);
        }
        put_k(i, j, v);
    });
    Mv.assign((size_t)c.MN.nnz(), 0.0);
    auto put_m = [&](int64_t i, int64_t j, double v) {
        const int64_t k = find_block(c.MN, i, j);
        if (k < 0) throw outside("M", i, j);
        Mv[c.MN.ptr[i] + k] += v;
    };
    for_each_csc(M, c.nP, c.nP, "M", [&](int64_t i, int64_t j, double v) {
        if (M_lower) {
            STK_CHECK(i >= j, STK_E_SHAPE, "M: entry above the diagonal with M_lower");
            if (i != j) put_m(j, i, v);
        }
        put_m(i, j, v);
    });
    Gv.assign(2 * (size_t)c.GN.nstore(), 0.0);
    GTv.assign(2 * (size_t)c.GTN.nstore(), 0.0);
    for_each_csc(G, c.nU, c.nP, "G", [&](int64_t i, int64_t j, double v) {
        const int64_t k = find_block(c.GN, i / 2, j), kt = find_block(c.GTN, j, i / 2);
        if (k < 0 || kt < 0) throw outside("G", i, j);
        if (c.fixed[i]) return;                          // rows of fixed dofs leave the solve
        Gv[2 * c.GN.pos(i / 2, k) + i % 2] += v;
        GTv[2 * c.GTN.pos(j, kt) + i % 2] += v;
    });
    STK_CHECK(Rhs, STK_E_ARG, "NULL Rhs");
    Fv.assign(Rhs, Rhs + c.nU);
}

void build_pressure_ic(Context &c, const std::vector<double> &M) {
    const Pattern &p = c.MN;
    const int64_t n = p.nrows;
    // greedy colouring of the vertex graph in natural order (3 colours on the 2-triangle split)
    std::vector<int32_t> colour(n, -1);
    int nc = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint64_t used = 0;
        for (int32_t k = p.ptr[i]; k < p.ptr[i + 1]; ++k)
            if (colour[p.col[k]] >= 0 && colour[p.col[k]] < 64) used |= 1ull << colour[p.col[k]];
        int ci = 0;
        while (used >> ci & 1ull) ++ci;
        STK_CHECK(ci < 64, STK_E_SHAPE, "pressure colouring needs more than 64 colours");
        colour[i] = ci;
        nc = std::max(nc, ci + 1);
    }
    std::vector<int32_t> order;
    order.reserve(n);
    c.pic.colour_off.assign(nc + 1, 0);
    for (int ci = 0; ci < nc; ++ci) {
        for (int64_t i = 0; i < n; ++i)
            if (colour[i] == ci) order.push_back((int32_t)i);
        c.pic.colour_off[ci + 1] = (int64_t)order.size();
    }
    std::vector<int64_t> pos(n);
    for (int64_t q = 0; q < n; ++q) pos[order[q]] = q;
    // up-looking IC(0): for row i (in colour order) and each earlier neighbour j,
    //   L_ij D_j = M_ij - sum_k L_ik D_k L_jk,   D_i = M_ii - sum_k L_ik^2 D_k,
    // k over earlier vertices in the pattern of both rows.  A pivot that is not positive restarts
    // the factorisation of M + alpha diag(M) with a larger alpha (the shift rule of limited-memory
    // incomplete Cholesky).
    std::vector<int32_t> lptr(n + 1, 0), lcol;
    std::vector<double> lval, D(n);
    std::vector<double> mark(n, 0.0);
    std::vector<uint8_t> has(n, 0);
    double alpha = 0.0;
    for (int attempt = 0;; ++attempt) {
        STK_CHECK(attempt < 40, STK_E_SHAPE, "incomplete Cholesky of MM: no positive pivots with any shift");
        lcol.clear();
        lval.clear();
        bool ok = true;
        for (int64_t q = 0; q < n && ok; ++q) {
            const int64_t i = order[q];
            lptr[q] = (int32_t)lcol.size();
            double dii = 0.0;
            std::vector<std::pair<int64_t, int32_t>> lower;   // (position, vertex) of earlier neighbours
            for (int32_t k = p.ptr[i]; k < p.ptr[i + 1]; ++k) {
                const int32_t j = p.col[k];
                if (j == i) dii = M[k] * (1.0 + alpha);
                else if (pos[j] < q) lower.push_back({pos[j], j});
            }
            std::sort(lower.begin(), lower.end());
            const size_t first = lcol.size();
            for (const auto &pj : lower) {
                const int32_t j = pj.second;
                double v = M[p.ptr[i] + find_block(p, i, j)];
                // sum over the entries k of row j (all earlier than j) that row i already holds;
                // row j is complete: its position precedes q, so lptr[pos_j + 1] is set
                for (int32_t t = lptr[pj.first]; t < lptr[pj.first + 1]; ++t) {
                    const int32_t kk = lcol[t];
                    if (has[kk]) v -= mark[kk] * D[kk] * lval[t];
                }
                const double lij = v / D[j];
                lcol.push_back(j);
                lval.push_back(lij);
                mark[j] = lij;
                has[j] = 1;
                dii -= lij * lij * D[j];
            }
            for (size_t t = first; t < lcol.size(); ++t) has[lcol[t]] = 0;
            if (!(dii > 1e-12 * std::fabs(M[p.ptr[i] + find_block(p, i, i)]))) ok = false;
            D[i] = dii;
        }
        if (ok) break;
        alpha = alpha > 0 ? 2.0 * alpha : 1e-3;
    }
    lptr[n] = (int32_t)lcol.size();
    // L^T by rows (position of the column vertex), for the backward solve
    std::vector<int32_t> uptr(n + 1, 0), ucol(lcol.size());
    std::vector<double> uval(lcol.size());
    for (size_t t = 0; t < lcol.size(); ++t) uptr[pos[lcol[t]] + 1]++;
    for (int64_t q = 0; q < n; ++q) uptr[q + 1] += uptr[q];
    std::vector<int32_t> fill(uptr.begin(), uptr.end() - 1);
    for (int64_t q = 0; q < n; ++q)
        for (int32_t t = lptr[q]; t < lptr[q + 1]; ++t) {
            const int64_t qc = pos[lcol[t]];
            ucol[fill[qc]] = order[q];
            uval[fill[qc]++] = lval[t];
        }
    std::vector<double> dinv(n);
    for (int64_t i = 0; i < n; ++i) dinv[i] = 1.0 / D[i];
    c.pic.order.upload(order, c.stream);
    c.pic.lptr.upload(lptr, c.stream);
    c.pic.lcol.upload(lcol, c.stream);
    c.pic.lval.upload(lval, c.stream);
    c.pic.uptr.upload(uptr, c.stream);
    c.pic.ucol.upload(ucol, c.stream);
    c.pic.uval.upload(uval, c.stream);
    c.pic.dinv.upload(dinv, c.stream);
    STK_HIP(hipStreamSynchronize(c.stream));
    c.pic.shift = alpha;
    c.pic.valid = true;
}

namespace {

// Local vertex lattice coordinates (i, j) of the 3 vertices of element type t in cell (I, J)
// (two triangles per cell).
inline void tri_vertices(int t, int I, int J, int vi[3], int vj[3]) {
    if (t == 0) {
        vi[0] = I;     vj[0] = J;
        vi[1] = I;     vj[1] = J + 1;
        vi[2] = I + 1; vj[2] = J;
    } else {
        vi[0] = I;     vj[0] = J + 1;
        vi[1] = I + 1; vj[1] = J + 1;
        vi[2] = I + 1; vj[2] = J;
    }
}

// Corner k = 0..3 of cell (I, J) in the order of Grid's quadrilaterals: (I,J), (I,J+1), (I+1,J+1),
// (I+1,J) as radial / azimuthal lattice offsets.
constexpr int kCornerDi[4] = {0, 0, 1, 1}, kCornerDj[4] = {0, 1, 1, 0};

void fill_level_geometry(MGLevel &L, hipStream_t s) {
    const int ncr = L.ncr, nct = L.nct, nr = ncr + 1;
    const int64_t ncell = (int64_t)ncr * nct, ncorner = (int64_t)nr * nct;
    L.nvert = L.split == 4 ? ncorner + ncell : ncorner;
    L.nel = (int64_t)L.split * ncell;
    L.h_e2v.resize(3 * L.nel);
    for (int64_t q = 0; q < ncell; ++q) {
        const int I = (int)(q % ncr), J = (int)(q / ncr);
        if (L.split == 4) {
            for (int t = 0; t < 4; ++t) {
                const int k0 = t, k1 = (t + 1) % 4;
                L.h_e2v[(4 * q + t) * 3 + 0] = (I + kCornerDi[k0]) + nr * ((J + kCornerDj[k0]) % nct);
                L.h_e2v[(4 * q + t) * 3 + 1] = (I + kCornerDi[k1]) + nr * ((J + kCornerDj[k1]) % nct);
                L.h_e2v[(4 * q + t) * 3 + 2] = (int32_t)(ncorner + q);
            }
            continue;
        }
        for (int t = 0; t < 2; ++t) {
            int vi[3], vj[3];
            tri_vertices(t, I, J, vi, vj);
            for (int a = 0; a < 3; ++a) L.h_e2v[(2 * q + t) * 3 + a] = vi[a] + nr * (vj[a] % nct);
        }
    }
    pattern_from_elements(L.pat, L.nvert, L.nvert, L.nel, L.h_e2v, 3, L.h_e2v, 3);
    L.pat.dptr.upload(L.pat.ptr, s);
    L.pat.dcol.upload(L.pat.col, s);
    if (L.nvert >= kSellMinRows) make_sliced(L.pat, s);
    std::vector<int32_t> perm;
    colour_elements(L.nel, 3, L.h_e2v, L.nvert, perm, L.col_off);
    std::vector<int32_t> e2v_p(3 * L.nel);
    std::vector<uint8_t> slot(9 * L.nel);
    for (int64_t p = 0; p < L.nel; ++p) {
        const int64_t e = perm[p];
        for (int a = 0; a < 3; ++a) {
            e2v_p[a * L.nel + p] = L.h_e2v[e * 3 + a];
            for (int b = 0; b < 3; ++b)
                slot[(a * 3 + b) * L.nel + p] =
                    (uint8_t)find_offset(L.pat, L.h_e2v[e * 3 + a], L.h_e2v[e * 3 + b]);
        }
    }
    L.perm.upload(perm, s);
    L.e2v.upload(e2v_p, s);
    L.slot.upload(slot, s);
    std::vector<double> cs(2 * L.nvert);
    std::vector<uint8_t> fixed(2 * L.nvert, 0);
    const double dth = 2.0 * M_PI / nct;
    for (int64_t v = 0; v < L.nvert; ++v) {
        // corners at angle dth j; centres half a cell further (radial dofs of wall corners fixed)
        const bool corner = v < ncorner;
        const int64_t q = v - ncorner;
        const int64_t i = corner ? v % nr : q % ncr;
        const double jj = corner ? (double)(v / nr) : (double)(q / ncr) + 0.5;
        cs[2 * v] = std::cos(dth * jj);
        cs[2 * v + 1] = std::sin(dth * jj);
        if (corner && (i == 0 || i == ncr)) fixed[2 * v + 1] = 1;
    }
    L.cs.upload(cs, s);
    L.fixed.upload(fixed, s);
    L.h_fixed = std::move(fixed);
}

// 2x2 rotation block R(d) = [[cos d, -sin d], [sin d, cos d]] scaled by w.
inline void rot_block(double w, double d, double *blk) {
    const double cd = std::cos(d), sd = std::sin(d);
    blk[0] = w * cd; blk[1] = -w * sd; blk[2] = w * sd; blk[3] = w * cd;
}

// Values of a block pattern given in CSR order, moved to the sliced positions (padding zero).
std::vector<double> sliced_values(const Pattern &p, const std::vector<double> &v, int bs) {
    std::vector<double> out((size_t)bs * p.nstore(), 0.0);
    for (int64_t r = 0; r < p.nrows; ++r)
        for (int32_t k = p.ptr[r]; k < p.ptr[r + 1]; ++k) {
            const int64_t q = p.pos(r, k - p.ptr[r]);
            for (int b = 0; b < bs; ++b) out[(size_t)bs * q + b] = v[(size_t)bs * k + b];
        }
    return out;
}

// Transfer operator as BSR 2x2 (fine node -> coarse vertex) plus its transpose.
struct Triplet { int32_t f, c; double blk[4]; };

void build_transfer(MGLevel &coarse, int64_t nfine, const std::vector<Triplet> &t,
                    const std::vector<uint8_t> &fine_fixed, const std::vector<uint8_t> &coarse_fixed,
                    hipStream_t s) {
    // P rows: fine nodes
    Pattern &P = coarse.P_pat;
    P.nrows = nfine; P.ncols = coarse.nvert;
    P.ptr.assign(nfine + 1, 0);
    for (const Triplet &x : t) P.ptr[x.f + 1]++;
    for (int64_t i = 0; i < nfine; ++i) P.ptr[i + 1] += P.ptr[i];
    P.col.resize(t.size());
    std::vector<double> Pv(4 * t.size());
    {
        std::vector<int32_t> pos(P.ptr.begin(), P.ptr.end() - 1);
        std::vector<const Triplet *> sorted(t.size());
        for (const Triplet &x : t) sorted[pos[x.f]++] = &x;
        for (int64_t i = 0; i < nfine; ++i)
            std::sort(sorted.begin() + P.ptr[i], sorted.begin() + P.ptr[i + 1],
                      [](const Triplet *a, const Triplet *b) { return a->c < b->c; });
        for (size_t k = 0; k < t.size(); ++k) {
            const Triplet &x = *sorted[k];
            P.col[k] = x.c;
            for (int q = 0; q < 4; ++q) {
                const int rr = q / 2, cc = q % 2;
                const bool zero = fine_fixed[2 * x.f + rr] || coarse_fixed[2 * x.c + cc];
                Pv[4 * k + q] = zero ? 0.0 : x.blk[q];
            }
        }
    }
    // R = P^T rows: coarse vertices
    Pattern &R = coarse.R_pat;
    R.nrows = coarse.nvert; R.ncols = nfine;
    R.ptr.assign(coarse.nvert + 1, 0);
    for (int32_t cidx : P.col) R.ptr[cidx + 1]++;
    for (int64_t i = 0; i < coarse.nvert; ++i) R.ptr[i + 1] += R.ptr[i];
    R.col.resize(P.col.size());
    std::vector<double> Rv(4 * P.col.size());
    std::vector<int32_t> pos(R.ptr.begin(), R.ptr.end() - 1);
    for (int64_t f = 0; f < nfine; ++f)
        for (int32_t k = P.ptr[f]; k < P.ptr[f + 1]; ++k) {
            const int32_t cidx = P.col[k];
            const int32_t dst = pos[cidx]++;
            R.col[dst] = (int32_t)f;
            // transpose of the 2x2 block
            Rv[4 * dst + 0] = Pv[4 * k + 0];
            Rv[4 * dst + 1] = Pv[4 * k + 2];
            Rv[4 * dst + 2] = Pv[4 * k + 1];
            Rv[4 * dst + 3] = Pv[4 * k + 3];
        }
    P.dptr.upload(P.ptr, s); P.dcol.upload(P.col, s);
    R.dptr.upload(R.ptr, s); R.dcol.upload(R.col, s);
    if (P.nrows >= kSellMinRows) { make_sliced(P, s); Pv = sliced_values(P, Pv, 4); }
    if (R.nrows >= kSellMinRows) { make_sliced(R, s); Rv = sliced_values(R, Rv, 4); }
    coarse.P64.upload(Pv, s);
    coarse.R64.upload(Rv, s);
    std::vector<float> Pf(Pv.begin(), Pv.end()), Rf(Rv.begin(), Rv.end());
    coarse.P32.upload(Pf, s);
    coarse.R32.upload(Rf, s);
}

}  // namespace

// Levels: mg[0] = P1 on the original mesh (vertices), mg[l+1] = 2x coarser P1 lattice of the same
// kind (two or four triangles per cell; the coarse triangles are unions of four fine ones).
// Transfer mg[0] -> P2 nodes interpolates midpoints from their edge ends; transfer
// mg[l+1] -> mg[l] interpolates the P1 field of the nested coarse triangles.  Both act on
// the rotated (u_theta, u_r) node frames, so every block is a rotation by the angle between
// the fine and the coarse node, and rows / columns of fixed (radial boundary) dofs are zero.
void build_mg_structure(Context &c) {
    c.mg.clear();
    if (!c.structured) return;
    int ncr = c.ncell_r, nct = c.ncell_t;
    const int split = c.mesh_split;
    auto nverts = [&](int64_t a, int64_t b) { return (a + 1) * b + (split == 4 ? a * b : 0); };
    {
        MGLevel L;
        L.ncr = ncr; L.nct = nct; L.split = split;
        c.mg.push_back(std::move(L));
    }
    // halve the lattice while it is above the size solved by a dense inverse
    while (ncr % 2 == 0 && nct % 2 == 0 && ncr / 2 >= 2 && nct / 2 >= 8 && nverts(ncr, nct) > kDenseMaxVerts) {
        ncr /= 2; nct /= 2;
        MGLevel L;
        L.ncr = ncr; L.nct = nct; L.split = split;
        c.mg.push_back(std::move(L));
    }
    for (MGLevel &L : c.mg) fill_level_geometry(L, c.stream);

    // P2 nodes <- vertices (level 0)
    {
        MGLevel &L0 = c.mg[0];
        std::vector<Triplet> t;
        std::vector<uint8_t> seen(c.nnod, 0);
        for (int64_t e = 0; e < c.nel; ++e) {
            const int32_t *n = &c.e2n[e * 6];
            const double *th = &c.el_th[e * 3];
            static const int ea[3] = {0, 1, 2}, eb[3] = {1, 2, 0};
            for (int a = 0; a < 3; ++a)
                if (!seen[n[a]]) {
                    seen[n[a]] = 1;
                    Triplet x{n[a], n[a], {1, 0, 0, 1}};
                    t.push_back(x);
                }
            for (int k = 0; k < 3; ++k) {
                const int32_t m = n[3 + k];
                if (seen[m]) continue;
                seen[m] = 1;
                const double half = 0.5 * (th[eb[k]] - th[ea[k]]);  // theta_m - theta_a
                Triplet xa{m, n[ea[k]], {}}, xb{m, n[eb[k]], {}};
                rot_block(0.5, half, xa.blk);
                rot_block(0.5, -half, xb.blk);
                t.push_back(xa);
                t.push_back(xb);
            }
        }
        build_transfer(L0, c.nnod, t, c.fixed, L0.h_fixed, c.stream);
    }
    // mg[l] <- mg[l+1]
    for (size_t l = 0; l + 1 < c.mg.size(); ++l) {
        MGLevel &F = c.mg[l], &C = c.mg[l + 1];
        const int nrf = F.ncr + 1, nrc = C.ncr + 1;
        const int64_t ncorner_f = (int64_t)nrf * F.nct, ncorner_c = (int64_t)nrc * C.nct;
        const double dthc = 2.0 * M_PI / C.nct;
        std::vector<Triplet> t;
        auto cv = [&](int ci, int cj) { return (int32_t)(ci + nrc * (cj % C.nct)); };
        auto ccentre = [&](int I, int J) { return (int32_t)(ncorner_c + I + (int64_t)C.ncr * J); };
        // fine angle minus coarse angle, both in units of the coarse cell width (unwrapped)
        auto add = [&](int64_t v, int32_t cvid, double fpos, double cpos, double w) {
            Triplet x{(int32_t)v, cvid, {}};
            rot_block(w, dthc * (fpos - cpos), x.blk);
            t.push_back(x);
        };
        for (int64_t v = 0; v < ncorner_f; ++v) {
            const int i = (int)(v % nrf), j = (int)(v / nrf);
            const int I = i / 2, J = j / 2;
            const double fpos = 0.5 * j;
            if (i % 2 == 0 && j % 2 == 0) add(v, cv(I, J), fpos, J, 1.0);
            else if (i % 2 == 1 && j % 2 == 0) { add(v, cv(I, J), fpos, J, 0.5); add(v, cv(I + 1, J), fpos, J, 0.5); }
            else if (i % 2 == 0 && j % 2 == 1) { add(v, cv(I, J), fpos, J, 0.5); add(v, cv(I, J + 1), fpos, J + 1, 0.5); }
            else if (F.split == 4) add(v, ccentre(I, J), fpos, J + 0.5, 1.0);   // the coarse cell's centre
            else { add(v, cv(I, J + 1), fpos, J + 1, 0.5); add(v, cv(I + 1, J), fpos, J, 0.5); }
        }
        // fine centres (four triangles per cell): the centre of fine cell (2I + a, 2J + b) is the
        // midpoint of the coarse centre and the coarse corner (I + a, J + b)
        for (int64_t v = ncorner_f; v < F.nvert; ++v) {
            const int64_t q = v - ncorner_f;
            const int fi = (int)(q % F.ncr), fj = (int)(q / F.ncr);
            const int I = fi / 2, J = fj / 2, a = fi % 2, b = fj % 2;
            const double fpos = 0.5 * (fj + 0.5);
            add(v, ccentre(I, J), fpos, J + 0.5, 0.5);
            add(v, cv(I + a, J + b), fpos, J + b, 0.5);
        }
        build_transfer(C, F.nvert, t, F.h_fixed, C.h_fixed, c.stream);
    }
    build_level_products(c);
    STK_HIP(hipStreamSynchronize(c.stream));
}

// Pattern of R A for every level from 1 on (R the restriction into it, A the operator of the
// next finer P1 level; level 0 has R K from the element assembly) and, per stored block of it,
// the pairs of stored positions (R block, A block) whose products sum there, in R's row order
// and then A's: the numeric product (k_pair_product) is a fixed-order sum, so it is
// deterministic.  The sliced layout follows R's.
void build_level_products(Context &c) {
    for (size_t k = 1; k < c.mg.size(); ++k) {
        MGLevel &C = c.mg[k], &F = c.mg[k - 1];
        const Pattern &R = C.R_pat, &A = F.pat;
        if (R.ncols != A.nrows) throw Error(STK_E_STATE, "restriction and level operator do not match");
        Pattern &Q = C.RA_pat;
        Q = Pattern{};
        Q.nrows = R.nrows;
        Q.ncols = A.ncols;
        Q.ptr.assign(Q.nrows + 1, 0);
        std::vector<int32_t> cols;
        for (int64_t u = 0; u < R.nrows; ++u) {
            cols.clear();
            for (int32_t i = R.ptr[u]; i < R.ptr[u + 1]; ++i) {
                const int32_t g = R.col[i];
                cols.insert(cols.end(), A.col.begin() + A.ptr[g], A.col.begin() + A.ptr[g + 1]);
            }
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            Q.col.insert(Q.col.end(), cols.begin(), cols.end());
            Q.ptr[u + 1] = (int32_t)Q.col.size();
        }
        Q.dptr.upload(Q.ptr, c.stream);
        Q.dcol.upload(Q.col, c.stream);
        if (R.sell) make_sliced(Q, c.stream);
        // pair lists, counted then filled, over the stored positions (padding keeps an empty list)
        const int64_t npos = Q.nstore();
        std::vector<int32_t> ptr(npos + 1, 0);
        auto each_pair = [&](auto &&fn) {
            for (int64_t u = 0; u < R.nrows; ++u)
                for (int32_t i = R.ptr[u]; i < R.ptr[u + 1]; ++i) {
                    const int32_t g = R.col[i];
                    const int64_t rpos = R.pos(u, i - R.ptr[u]);
                    for (int32_t j = A.ptr[g]; j < A.ptr[g + 1]; ++j)
                        fn(Q.pos(u, find_offset(Q, (int32_t)u, A.col[j])), rpos, A.pos(g, j - A.ptr[g]));
                }
        };
        each_pair([&](int64_t q, int64_t, int64_t) { ptr[q + 1]++; });
        for (int64_t q = 0; q < npos; ++q) ptr[q + 1] += ptr[q];
        std::vector<int32_t> pair(2 * (size_t)ptr[npos]);
        std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
        each_pair([&](int64_t q, int64_t rpos, int64_t apos) {
            const int64_t at = fill[q]++;
            pair[2 * at] = (int32_t)rpos;
            pair[2 * at + 1] = (int32_t)apos;
        });
        C.ra_ptr.upload(ptr, c.stream);
        C.ra_pair.upload(pair, c.stream);
        STK_HIP(hipStreamSynchronize(c.stream));   // the host vectors go out of scope
    }
}

}  // namespace stk
