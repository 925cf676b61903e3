// Numeric multifrontal factorisation on gfx950.
//
// Fronts are processed level by level (height above the leaves), all fronts
// of a level and all frequencies of a batch in the same launches:
//   1. assemble: every entry of a front is written once, as the sum of the
//      children's Schur complements (pulled through child->parent index maps);
//   2. the original matrix entries owned by the level are added (each
//      destination is unique, so no atomics);
//   3. partial LU of the first p columns with row pivoting restricted to the
//      p fully summed rows (the front's structure never changes):
//      - fronts with m <= small_max: one workgroup, front held in LDS;
//      - larger fronts: blocked right-looking LU, per panel step a panel
//        kernel (pivot search, row swaps over the whole front, scaling),
//        a TRSM of the U row block, and an MFMA trailing update of the
//        remaining front (which leaves the Schur complement in F22).
// Factors stay in place in the front arena (L11\U11 and L21 in the first p
// columns, U12 in the first p rows); the pivot swaps go to piv[pos0 + j].
#include <algorithm>

#include "engine.hpp"
#include "mfma.hpp"
#include "prof.hpp"

namespace gfdfd {


// ---------------------------------------------------------------------------
// DevPlan
// ---------------------------------------------------------------------------
// Small subtrees: a front belongs when its order is at most the plan's
// small_max and all its children belong (leaves included), so a whole subtree
// is factored by the persistent launch and the level path never waits on it.
// Fronts are in postorder, so each maximal subtree is a contiguous range
// ending at its root.
void DevPlan::build_subtrees(const Plan &pl) {
    const int nfr = (int)pl.fronts.size();
    std::vector<char> in(nfr, 0);
    for (int f = 0; f < nfr; f++) {   // postorder: children first
        const Front &F = pl.fronts[f];
        bool ok = pl.grid && F.m <= pl.small_max && F.m <= kSubMaxM && F.nchild <= kAsmMaxChildren;
        for (int c = 0; c < F.nchild && ok; c++) ok = in[pl.children[F.child_off + c]] != 0;
        in[f] = ok;
    }
    std::vector<int> list, slot(nfr, -1), first, last, lo(nfr);
    for (int f = 0; f < nfr; f++) {   // first (lowest) front of f's subtree
        const Front &F = pl.fronts[f];
        lo[f] = f;
        for (int c = 0; c < F.nchild; c++) lo[f] = std::min(lo[f], lo[pl.children[F.child_off + c]]);
    }
    for (int f = 0; f < nfr; f++)
        if (in[f]) {
            slot[f] = (int)list.size();
            list.push_back(f);
            const int par = pl.fronts[f].parent;
            if (par < 0 || !in[par]) {
                for (int g = lo[f]; g <= f; g++)
                    if (!in[g]) throw std::logic_error("small subtree is not a contiguous postorder range");
                first.push_back(lo[f]);
                last.push_back(f);
            }
        }
    nsub = (int)list.size();
    nroots = (int)first.size();
    // original entries: the subtrees' by slot (local offsets), the rest by level
    std::vector<std::vector<std::pair<int64_t, int>>> per_slot(nsub);
    std::vector<std::vector<std::pair<int64_t, int64_t>>> rest(pl.nlevels);
    std::vector<std::pair<int64_t, int>> starts;   // buf_off -> front
    for (int f = 0; f < nfr; f++) starts.push_back({pl.fronts[f].buf_off, f});
    std::sort(starts.begin(), starts.end());
    for (size_t e = 0; e < pl.orig_q.size(); e++) {
        const int64_t d = pl.orig_dst[e];
        auto it = std::upper_bound(starts.begin(), starts.end(), std::make_pair(d, INT32_MAX));
        if (it == starts.begin()) throw std::logic_error("original entry before the first front");
        const int f = std::prev(it)->second;
        const Front &F = pl.fronts[f];
        if (d - F.buf_off >= (int64_t)F.m * F.m) throw std::logic_error("original entry outside its front");
        if (slot[f] >= 0) per_slot[slot[f]].push_back({pl.orig_q[e], (int)(d - F.buf_off)});
        else rest[F.level].push_back({pl.orig_q[e], d});
    }
    std::vector<int64_t> optr(nsub + 1, 0), oq;
    std::vector<int> ooff;
    for (int k = 0; k < nsub; k++) {
        for (auto &e : per_slot[k]) { oq.push_back(e.first); ooff.push_back(e.second); }
        optr[k + 1] = (int64_t)oq.size();
    }
    std::vector<int64_t> rq, rd;
    for (int l = 0; l < pl.nlevels; l++) {
        const int64_t b = (int64_t)rq.size();
        for (auto &e : rest[l]) { rq.push_back(e.first); rd.push_back(e.second); }
        orig_rest_range.push_back({b, (int64_t)rq.size()});
    }
    root_first = dev_upload(first);
    root_last = dev_upload(last);
    slot_of = dev_upload(slot);
    sub_optr = dev_upload(optr);
    sub_oq = dev_upload(oq);
    sub_ooff = dev_upload(ooff);
    orig_q_rest = dev_upload(rq);
    orig_dst_rest = dev_upload(rd);
    bytes += (2 * first.size() + slot.size() + ooff.size()) * 4 + (optr.size() + oq.size() + rq.size() + rd.size()) * 8;
    for (int f : list) {
        const Front &F = pl.fronts[f];
        sub_mmax = std::max(sub_mmax, F.m);
        for (int k = 0; k < F.p; k++) {
            const double t = F.m - k - 1;
            sub_flops += 4.0 * (2.0 * t * t + t);
        }
        sub_elems += (double)F.m * F.m;
    }
    // the level path's lists of the remaining fronts
    for (int l = 0; l < pl.nlevels; l++) {
        const LevelPlan &L = pl.levels[l];
        std::vector<int> a, sm;
        int mm = 0, smm = 0;
        double el = 0, sfl = 0, sel = 0;
        for (int f : L.small)
            if (slot[f] < 0) {
                const Front &F = pl.fronts[f];
                a.push_back(f);
                sm.push_back(f);
                mm = std::max(mm, F.m);
                smm = std::max(smm, F.m);
                el += (double)F.m * F.m;
                sel += 2.0 * F.m * F.m;
                for (int k = 0; k < F.p; k++) {
                    const double t = F.m - k - 1;
                    sfl += 4.0 * (2.0 * t * t + t);
                }
            }
        for (int f : L.large) {
            a.push_back(f);
            mm = std::max(mm, pl.fronts[f].m);
            el += (double)pl.fronts[f].m * pl.fronts[f].m;
        }
        lv_asm_rest.push_back(dev_upload(a));
        lv_small_rest.push_back(dev_upload(sm));
        n_asm_rest.push_back((int)a.size());
        n_small_rest.push_back((int)sm.size());
        mmax_asm_rest.push_back(mm);
        small_mmax_rest.push_back(smm);
        asm_rest_elems.push_back(el);
        small_rest_flops.push_back(sfl);
        small_rest_elems.push_back(sel);
    }
}

DevPlan::DevPlan(std::shared_ptr<const Plan> p) : host(std::move(p)) {
    const Plan &pl = *host;
    std::vector<FrontDev> fd(pl.fronts.size());
    for (size_t f = 0; f < pl.fronts.size(); f++) {
        const Front &F = pl.fronts[f];
        fd[f] = {F.p, F.m, F.pos0, F.nchild, F.child_off, F.level, F.buf_off, F.fv_off, F.idx_off, F.inv_off};
    }
    fronts = dev_upload(fd);
    children = dev_upload(pl.children);
    idx = dev_upload(pl.idx);
    inv = dev_upload(pl.inv);
    perm = dev_upload(pl.perm);
    iperm = dev_upload(pl.iperm);
    orig_q = dev_upload(pl.orig_q);
    orig_dst = dev_upload(pl.orig_dst);
    colptr = dev_upload(pl.colptr);
    rowind = dev_upload(pl.rowind);
    bytes += fd.size() * sizeof(FrontDev) + (pl.children.size() + pl.idx.size() + pl.inv.size() +
                                            2 * pl.perm.size() + pl.colptr.size() + pl.rowind.size()) * 4 +
             (pl.orig_q.size() + pl.orig_dst.size()) * 8;
    // gather sources of every front row before pivoting, at the front's work rows:
    // [X position or -1, W row of a child's contribution or -1, a second one or -1]
    {
        std::vector<int> g((size_t)pl.fv_rows * 3, -1);
        for (const Front &F : pl.fronts)
            for (int i = 0; i < F.m; i++) {
                int *r = &g[(size_t)(F.fv_off + i) * 3];
                if (i < F.p) r[0] = F.pos0 + i;
                int k = 1;
                for (int c = 0; c < F.nchild; c++) {
                    const Front &C = pl.fronts[pl.children[F.child_off + c]];
                    const int a = pl.inv[C.inv_off + i];
                    if (a < 0) continue;
                    if (k == 3) throw std::logic_error("front row with more than two child contributions");
                    r[k++] = (int)(C.fv_off + C.p + a);
                }
            }
        gsrc_id = dev_upload(g);
        bytes += g.size() * 4;
    }
    // CSR-of-M -> CSC map, valid when the pattern is symmetric (grid plans)
    if (pl.grid) {
        std::vector<int64_t> t(pl.nnz);
        std::vector<int> fill(pl.n, 0);
        for (int c = 0; c < pl.n; c++)
            for (int q = pl.colptr[c]; q < pl.colptr[c + 1]; q++) {
                int r = pl.rowind[q];
                // row r of M: its entries in column order are the rows of CSC column r (symmetric pattern)
                t[pl.colptr[r] + fill[r]++] = q;
            }
        tq = dev_upload(t);
        bytes += t.size() * 8;
    }
    for (const LevelPlan &L : pl.levels) {
        std::vector<int> all(L.small);
        all.insert(all.end(), L.large.begin(), L.large.end());
        std::vector<FrontDev> lf;
        int haskids = 0;
        for (int f : all) {
            lf.push_back(fd[f]);
            haskids |= pl.fronts[f].nchild > 0;
        }
        lv_children.push_back(haskids);
        lv_fr.push_back(dev_upload(lf));
        lv_small.push_back(dev_upload(L.small));
        lv_large.push_back(dev_upload(L.large));
        lv_all.push_back(dev_upload(all));
        n_small.push_back((int)L.small.size());
        n_large.push_back((int)L.large.size());
        n_all.push_back((int)all.size());
        mmax_all.push_back(std::max(L.small_mmax, L.large_mmax));
        int pm = 0;
        for (int f : all) pm = std::max(pm, pl.fronts[f].p);
        pmax_all.push_back(pm);
    }
    // tile lists of every blocked step
    std::vector<TileDev> tl;
    std::vector<Tile> tmp;
    steps.resize(pl.levels.size());
    for (size_t l = 0; l < pl.levels.size(); l++) {
        const LevelPlan &L = pl.levels[l];
        for (int k0 = 0; k0 < L.large_pmax; k0 += pl.nb) {
            step_tiles(pl, L, k0, kTile, tmp);
            steps[l].push_back({(int64_t)tl.size(), (int)tmp.size()});
            for (const Tile &t : tmp) tl.push_back({t.front, t.ti, t.tj});
        }
    }
    // the two-panel schedule: per 64-column block b0 of a level, the thin update
    // (rows from b0 + 32, the second panel's columns: fronts with p > b0 + 32)
    // and the depth-min(64, p - b0) trailing update
    if (pl.nb != 32) throw std::logic_error("the blocked path assumes 32-column panels");
    deep_steps.resize(pl.levels.size());
    for (size_t l = 0; l < pl.levels.size(); l++) {
        const LevelPlan &L = pl.levels[l];
        for (int b0 = 0; b0 < L.large_pmax; b0 += 64) {
            DeepStep d;
            d.thin_off = (int64_t)tl.size();
            for (int li = 0; li < (int)L.large.size(); li++) {
                const Front &F = pl.fronts[L.large[li]];
                if (F.p <= b0 + 32) continue;
                const int kb2 = std::min(32, F.p - b0 - 32), t = F.m - b0 - 32;
                for (int ti = 0; ti < (t + kTile - 1) / kTile; ti++) tl.push_back({li, ti, 0});
                d.thin_flops += 8.0 * t * kb2 * 32;
                d.thin_elems += 2.0 * t * kb2 + t * 32.0 + 32.0 * kb2;
                d.trsm_thin_flops += 8.0 * 0.5 * 32 * 32 * kb2;
            }
            d.thin_cnt = (int)(tl.size() - d.thin_off);
            d.deep_off = (int64_t)tl.size();
            for (int li = 0; li < (int)L.large.size(); li++) {
                const Front &F = pl.fronts[L.large[li]];
                if (F.p <= b0) continue;
                const int kd = std::min(64, F.p - b0), t = F.m - b0 - kd;
                for (int j = b0; j < b0 + kd; j++) {
                    const int pe = j < b0 + 32 ? std::min(b0 + 32, b0 + kd) : b0 + kd;   // the panel's end
                    (j < b0 + 32 ? d.panel1_flops : d.panel2_flops) +=
                        8.0 * (F.m - j - 1) * (pe - j - 1) + 6.0 * (F.m - j - 1);
                }
                if (t <= 0) continue;
                const int nt = (t + kTile - 1) / kTile;
                for (int tj = 0; tj < nt; tj++)
                    for (int ti = 0; ti < nt; ti++) tl.push_back({li, ti, tj});
                d.deep_flops += 8.0 * t * t * kd;
                d.deep_elems += 2.0 * t * kd + 2.0 * t * t;
                d.trsm_deep_flops += 8.0 * 0.5 * kd * kd * t;
            }
            d.deep_cnt = (int)(tl.size() - d.deep_off);
            deep_steps[l].push_back(d);
        }
    }
    tiles = dev_upload(tl);
    bytes += tl.size() * sizeof(TileDev);
    for (const Front &F : pl.fronts) max_children = std::max(max_children, F.nchild);
    // single-block levels: every front of the level has at most one pivot block
    // and is stored in the one-product solve format (see k_diag_inv)
    for (size_t l = 0; l < pl.levels.size(); l++)
        single.push_back(pmax_all[l] <= 32 && mmax_all[l] <= kSingleMmax && max_children <= 2 ? 1 : 0);
    lv_single = dev_upload(single);
    // inverse squares for the levels of fewer than kVecSplitFronts multi-block
    // fronts whose pivot blocks fit the inversion kernel's LDS
    {
        std::vector<int> lst;
        std::vector<int64_t> loff;
        vinv.assign(pl.levels.size(), 0);
        vinv_off.assign(pl.levels.size(), nullptr);
        for (size_t l = 0; l < pl.levels.size(); l++) {
            if (single[l] || n_all[l] >= kVecSplitFronts || pmax_all[l] > kVinvPmax || n_all[l] == 0) continue;
            vinv[l] = 1;
            const LevelPlan &L = pl.levels[l];
            std::vector<int> all(L.small);
            all.insert(all.end(), L.large.begin(), L.large.end());
            std::vector<int64_t> off;
            for (int f : all) {
                const int p = pl.fronts[f].p, c = pl.fronts[f].m - p;
                off.push_back(vinv_elems);
                lst.push_back(f);
                loff.push_back(vinv_elems);
                // the square, then L21 inv(L11) (c x p) and inv(U11) U12 (p x c)
                vinv_elems += (int64_t)p * p + 2 * (int64_t)p * c;
                vinv_pmax = std::max(vinv_pmax, p);
                vinv_rect_tiles = std::max(vinv_rect_tiles, 2 * ((p + 63) / 64) * ((c + 63) / 64));
            }
            vinv_off[l] = dev_upload(off);
        }
        vinv_n = (int)lst.size();
        vinv_list = dev_upload(lst);
        vinv_list_off = dev_upload(loff);
        bytes += lst.size() * 12;
    }
    build_subtrees(pl);
    counts.resize(pl.levels.size());
    for (size_t l = 0; l < pl.levels.size(); l++) {
        const LevelPlan &L = pl.levels[l];
        LevelCounts &c = counts[l];
        auto each = [&](const std::vector<int> &v, auto fn) { for (int f : v) fn(pl.fronts[f]); };
        auto tri = [&](const Front &F) {
            double p = F.p, m = F.m;
            c.tri_elems += 0.5 * p * p + p * (m - p);
            c.sum_m += m;
            c.asm_elems += m * m;
        };
        each(L.small, tri);
        each(L.large, tri);
        each(L.small, [&](const Front &F) {
            for (int k = 0; k < F.p; k++) {
                double t = F.m - k - 1;
                c.small_flops += 4.0 * (2.0 * t * t + t);
            }
            c.small_elems += 2.0 * F.m * F.m;
        });
        each(L.small, [&](const Front &F) { c.contrib_elems += (double)F.p * (F.m - F.p); c.sum_p += F.p; });
        each(L.large, [&](const Front &F) { c.contrib_elems += (double)F.p * (F.m - F.p); c.sum_p += F.p; });
        int lpmax = 0;
        for (int f : L.small) lpmax = std::max(lpmax, pl.fronts[f].p);
        for (int f : L.large) lpmax = std::max(lpmax, pl.fronts[f].p);
        for (int b0 = 0; b0 < lpmax; b0 += 32) {
            double fe = 0, be = 0;
            auto blk = [&](const Front &F) {
                if (F.p <= b0) return;
                const double kb = std::min(32, F.p - b0);
                fe += 0.5 * kb * kb + kb * (F.m - b0 - kb);
                be += 0.5 * kb * kb + kb * b0;
                diag_elems += kb * kb;
            };
            each(L.small, blk);
            each(L.large, blk);
            c.fwd_blk_elems.push_back(fe);
            c.bwd_blk_elems.push_back(be);
        }
        for (int k0 = 0; k0 < L.large_pmax; k0 += pl.nb) {
            double gf = 0, ge = 0, pf = 0, tf = 0;
            each(L.large, [&](const Front &F) {
                if (F.p <= k0) return;
                double kb = std::min(pl.nb, F.p - k0), t = F.m - k0 - kb;
                gf += 8.0 * t * t * kb;
                ge += 2.0 * t * kb + 2.0 * t * t;
                for (int j = k0; j < k0 + kb; j++) pf += 8.0 * (F.m - j - 1) * (k0 + kb - j - 1) + 6.0 * (F.m - j - 1);
                tf += 8.0 * 0.5 * kb * kb * t;
            });
            c.gemm_flops.push_back(gf);
            c.gemm_elems.push_back(ge);
            c.panel_flops.push_back(pf);
            c.trsm_flops.push_back(tf);
        }
    }
}

DevPlan::~DevPlan() {
    dev_free(fronts); dev_free(children); dev_free(idx); dev_free(inv); dev_free(perm); dev_free(iperm);
    dev_free(orig_q); dev_free(orig_dst); dev_free(colptr); dev_free(rowind); dev_free(tq); dev_free(tiles);
    dev_free(lv_single);
    dev_free(gsrc_id);
    for (auto p : vinv_off) dev_free(p);
    dev_free(vinv_list);
    dev_free(vinv_list_off);
    dev_free(root_first); dev_free(root_last); dev_free(slot_of); dev_free(sub_optr); dev_free(sub_oq);
    dev_free(sub_ooff); dev_free(orig_q_rest); dev_free(orig_dst_rest);
    for (auto p : lv_asm_rest) dev_free(p);
    for (auto p : lv_small_rest) dev_free(p);
    for (auto p : lv_fr) dev_free(p);
    for (auto p : lv_small) dev_free(p);
    for (auto p : lv_large) dev_free(p);
    for (auto p : lv_all) dev_free(p);
}

template <typename T>
Factors<T>::~Factors() {
    if (graph) (void)hipGraphExecDestroy(graph);
    dev_free(arena);
    dev_free(piv);
    dev_free(rowperm);
    dev_free(sing);
    dev_free(gsrc);
    dev_free(sched);
    dev_free(vinv);
    if (own_vals) dev_free(vals);
}
template struct Factors<double>;
template struct Factors<float>;

// ---------------------------------------------------------------------------
// 1+2: front assembly
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_front_assemble(cx<T> *__restrict__ arena, int64_t fstride,
                                                        const FrontDev *__restrict__ fronts,
                                                        const int *__restrict__ list, const int *__restrict__ children,
                                                        const int *__restrict__ inv) {
    extern __shared__ int cmap[];   // [nchild][m]: child row of each front row, or -1
    __shared__ int64_t cbase[kAsmMaxChildren];
    __shared__ int cstride[kAsmMaxChildren];
    const FrontDev F = fronts[list[blockIdx.x]];
    cx<T> *A = arena + blockIdx.z * fstride;
    cx<T> *G = A + F.buf_off;
    const int m = F.m, nc = F.nchild;
    // the children's row maps and block origins, staged once: each entry below
    // then waits on one value load only
    if (threadIdx.x < nc) {
        const FrontDev C = fronts[children[F.child_off + threadIdx.x]];
        cbase[threadIdx.x] = C.buf_off + C.p + (int64_t)C.p * C.m;   // contribution block origin
        cstride[threadIdx.x] = C.m;
    }
    for (int c = 0; c < nc; c++) {
        const int io = fronts[children[F.child_off + c]].inv_off;
        for (int r = threadIdx.x; r < m; r += blockDim.x) cmap[c * m + r] = inv[io + r];
    }
    __syncthreads();
    const int64_t total = (int64_t)m * m;
    const int64_t step = (int64_t)gridDim.y * blockDim.x;
    if (nc <= 2 && total < (int64_t(1) << 31)) {
        // up to two children (the grid plans' fronts): each entry's child
        // values are loaded unconditionally (an entry a child does not cover
        // reads the child's block origin and is dropped), four entries at a
        // time, so eight loads are in flight before the first add
        // (entry indices in 32 bits)
        constexpr int U = 4;
        const int tot = (int)total, stp = (int)step;
        for (int e0 = blockIdx.y * blockDim.x + threadIdx.x; e0 < tot; e0 += U * stp) {
            cx<T> w[U][2];
            bool ok[U][2];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = e0 + u * stp;
                const int ec = e < tot ? e : 0;
                const int s = ec / m, r = ec - s * m;
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const int a = c < nc ? cmap[c * m + r] : -1, b = c < nc ? cmap[c * m + s] : -1;
                    ok[u][c] = a >= 0 && b >= 0;
                    w[u][c] = A[c < nc ? cbase[c] + (ok[u][c] ? a + (int64_t)b * cstride[c] : 0) : F.buf_off];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = e0 + u * stp;
                if (e >= tot) break;
                cx<T> v = {T(0), T(0)};
                if (ok[u][0]) v = w[u][0];
                if (ok[u][1]) v = cadd(v, w[u][1]);
                G[e] = v;
            }
        }
        return;
    }
    // unrolled: several entries' value loads in flight at once
#pragma unroll 4
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < total; e += step) {
        const int s = (int)(e / m), r = (int)(e - (int64_t)s * m);
        cx<T> v = {T(0), T(0)};
        for (int c = 0; c < nc; c++) {
            const int a = cmap[c * m + r], b = cmap[c * m + s];
            if (a >= 0 && b >= 0) v = cadd(v, A[cbase[c] + a + (int64_t)b * cstride[c]]);
        }
        G[e] = v;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_scatter_orig(cx<T> *__restrict__ arena, int64_t fstride,
                                                      const cx<double> *__restrict__ vals, int64_t vstride,
                                                      const int64_t *__restrict__ q, const int64_t *__restrict__ dst,
                                                      int64_t count) {
    cx<T> *A = arena + blockIdx.y * fstride;
    const cx<double> *V = vals + blockIdx.y * vstride;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x) {
        cx<double> v = V[q[e]];
        cx<T> &d = A[dst[e]];
        d.x += (T)v.x;
        d.y += (T)v.y;
    }
}

// ---------------------------------------------------------------------------
// pivot search helper: max cabs1 with the smallest row on ties
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void better(T &best, int &bi, T ob, int oi) {
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
}

template <typename T>
__device__ int block_argmax(T best, int bi, T *sb, int *si) {
    // wave reduction then across waves through LDS
    for (int off = 32; off > 0; off >>= 1) {
        T ob = __shfl_xor(best, off);
        int oi = __shfl_xor(bi, off);
        better(best, bi, ob, oi);
    }
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) { sb[w] = best; si[w] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < nw; k++) better(sb[0], si[0], sb[k], si[k]);
    }
    __syncthreads();
    return si[0];
}

// ---------------------------------------------------------------------------
// 3a: small fronts, one workgroup, front in LDS
// ---------------------------------------------------------------------------
// Partial LU of an m x m front held column-major in LDS (p pivots, row
// pivoting within the p fully-summed rows); the swap of step j goes to piv[j].
// All threads of the workgroup call it (at least one wave); sp holds kLuNb
// ints of LDS.  Blocked by kLuNb pivots: wave 0 alone factors a panel of kLuNb
// columns -- pivot search, row swap within the panel, scaling and the rank-1
// update of the panel's columns, lanes over rows, with no barrier (a wave's
// LDS operations complete in order) -- then the workgroup applies the panel's
// swaps to the other columns, forms the panel's rows of U12 by forward
// substitution with the panel's unit lower block, and makes one trailing update
// A22 -= L21 U12 of depth kLuNb with 2 x 2 output tiles per thread, so the
// trailing block passes through LDS once per panel instead of once per pivot.
// Complex products use the explicit fused multiply-adds (cmul_f, cmsub_f): the
// persistent subtree kernel and the level path inline this function and must
// round alike.
constexpr int kLuNb = 8;
template <typename T>
__device__ void small_lu_lds(cx<T> *S, int m, int p, int *__restrict__ piv, int *__restrict__ sing, int *sp) {
    const int tid = threadIdx.x, bs = blockDim.x, lane = tid & 63;
    for (int jb = 0; jb < p; jb += kLuNb) {
        const int je = min(p, jb + kLuNb);   // panel columns [jb, je)
        if (tid < 64) {
            for (int j = jb; j < je; j++) {
                // pivot: largest |.|_1 over rows [j, p), ties to the lower row
                T best = T(-1);
                int bi = j;
                for (int r = j + lane; r < p; r += 64) better(best, bi, cabs1(S[r + j * m]), r);
                for (int off = 32; off > 0; off >>= 1) {
                    const T ob = __shfl_xor(best, off);
                    const int oi = __shfl_xor(bi, off);
                    better(best, bi, ob, oi);
                }
                if (lane == 0) {
                    sp[j - jb] = bi;
                    piv[j] = bi;
                    if (best == T(0)) atomicOr(sing, 1);
                }
                // swap rows j and bi in the panel's columns
                if (bi != j && jb + lane < je) {
                    const int c = jb + lane;
                    const cx<T> t = S[j + c * m];
                    S[j + c * m] = S[bi + c * m];
                    S[bi + c * m] = t;
                }
                const cx<T> d = crecip(S[j + j * m]);
                for (int r = j + 1 + lane; r < m; r += 64) {
                    const cx<T> l = cmul_f(S[r + j * m], d);
                    S[r + j * m] = l;
                    for (int c = j + 1; c < je; c++) S[r + c * m] = cmsub_f(S[r + c * m], l, S[j + c * m]);
                }
            }
        }
        __syncthreads();
        // the panel's swaps, in order, on the columns outside it
        for (int c = tid; c < m; c += bs) {
            if (c >= jb && c < je) continue;
            for (int j = jb; j < je; j++) {
                const int r = sp[j - jb];
                if (r != j) {
                    const cx<T> t = S[j + c * m];
                    S[j + c * m] = S[r + c * m];
                    S[r + c * m] = t;
                }
            }
        }
        __syncthreads();
        if (je >= m) break;
        // U12 rows [jb, je) of columns [je, m): one thread per column, rows in order
        for (int c = je + tid; c < m; c += bs)
            for (int i = jb + 1; i < je; i++) {
                cx<T> a = S[i + c * m];
                for (int t = jb; t < i; t++) a = cmsub_f(a, S[i + t * m], S[t + c * m]);
                S[i + c * m] = a;
            }
        __syncthreads();
        // trailing update of rows and columns [je, m): 2 x 2 tiles, the tile's
        // rows fastest across threads
        const int q = m - je, qt = (q + 1) >> 1;
        for (int tt = tid; tt < qt * qt; tt += bs) {
            const int r0 = je + 2 * (tt % qt), c0 = je + 2 * (tt / qt);
            const bool r1 = r0 + 1 < m, c1 = c0 + 1 < m;
            cx<T> a00 = S[r0 + c0 * m], a10 = r1 ? S[r0 + 1 + c0 * m] : cx<T>{T(0), T(0)};
            cx<T> a01 = c1 ? S[r0 + (c0 + 1) * m] : cx<T>{T(0), T(0)};
            cx<T> a11 = (r1 && c1) ? S[r0 + 1 + (c0 + 1) * m] : cx<T>{T(0), T(0)};
#pragma unroll 2
            for (int t = jb; t < je; t++) {
                const cx<T> l0 = S[r0 + t * m], l1 = r1 ? S[r0 + 1 + t * m] : cx<T>{T(0), T(0)};
                const cx<T> u0 = S[t + c0 * m], u1 = c1 ? S[t + (c0 + 1) * m] : cx<T>{T(0), T(0)};
                a00 = cmsub_f(a00, l0, u0);
                a10 = cmsub_f(a10, l1, u0);
                a01 = cmsub_f(a01, l0, u1);
                a11 = cmsub_f(a11, l1, u1);
            }
            S[r0 + c0 * m] = a00;
            if (r1) S[r0 + 1 + c0 * m] = a10;
            if (c1) S[r0 + (c0 + 1) * m] = a01;
            if (r1 && c1) S[r0 + 1 + (c0 + 1) * m] = a11;
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_small_lu(cx<T> *__restrict__ arena, int64_t fstride,
                                                  const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                  int *__restrict__ piv, int n, int *__restrict__ sing) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *S = reinterpret_cast<cx<T> *>(smem_raw);
    __shared__ int sp[kLuNb];   // the current panel's pivot rows
    const FrontDev F = fronts[list[blockIdx.x]];
    const int fq = blockIdx.y;
    cx<T> *G = arena + fq * fstride + F.buf_off;
    const int m = F.m, mm = m * m, tid = threadIdx.x, bs = blockDim.x;
    // FP32: the front staged with all of a thread's loads in flight (m <= 64: at
    // most 16 entries per thread at 256 threads; entries past the front read its
    // first).  FP64 keeps one round per 4 entries, measured faster at its
    // register budget.
    if constexpr (sizeof(T) == 8) {
#pragma unroll 4
        for (int e = tid; e < mm; e += bs) S[e] = G[e];
    } else {
        for (int e0 = 0; e0 < mm; e0 += 16 * bs) {
            cx<T> v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) {
                const int e = e0 + tid + u * bs;
                v[u] = G[e < mm ? e : 0];
            }
#pragma unroll
            for (int u = 0; u < 16; u++) {
                const int e = e0 + tid + u * bs;
                if (e < mm) S[e] = v[u];
            }
        }
    }
    __syncthreads();
    small_lu_lds<T>(S, m, F.p, piv + (int64_t)fq * n + F.pos0, sing + fq, sp);
    for (int e = tid; e < mm; e += bs) G[e] = S[e];
}

// ---------------------------------------------------------------------------
// 3a': the small subtrees in one persistent launch.  A task is one maximal
// small subtree (fronts first..last, children before parents) for one
// frequency; workgroups take tasks from a global counter until none is left.
// A workgroup factors its subtree's fronts in order: each front is assembled
// in LDS from its children's contribution blocks -- written by this same
// workgroup earlier in the task, so a workgroup barrier is all the
// synchronisation there is -- and its original entries, factored there and
// written to the arena.  No workgroup ever waits for another.  Same
// arithmetic, in the same order, as k_front_assemble + k_scatter_orig +
// k_small_lu.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_subtree_factor(cx<T> *__restrict__ arena, int64_t fstride,
                                                        const FrontDev *__restrict__ fronts,
                                                        const int *__restrict__ root_first,
                                                        const int *__restrict__ root_last,
                                                        const int *__restrict__ slot_of,
                                                        const int *__restrict__ children, const int *__restrict__ inv,
                                                        const int64_t *__restrict__ optr,
                                                        const int64_t *__restrict__ oq, const int *__restrict__ ooff,
                                                        const cx<double> *__restrict__ vals, int64_t vstride,
                                                        int nroots, int nf, int *__restrict__ piv, int n,
                                                        int *__restrict__ sing, int *__restrict__ counter) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *S = reinterpret_cast<cx<T> *>(smem_raw);
    __shared__ int cmap[kAsmMaxChildren][kSubMaxM];
    __shared__ int64_t cbase[kAsmMaxChildren];
    __shared__ int cstride[kAsmMaxChildren];
    __shared__ int task;
    __shared__ int sp[kLuNb];   // the current panel's pivot rows
    const int tid = threadIdx.x, ntask = nroots * nf;
    for (;;) {
        if (tid == 0) task = atomicAdd(counter, 1);
        __syncthreads();
        const int t = task;
        if (t >= ntask) return;
        const int r = t / nf, fq = t - r * nf;
        cx<T> *A = arena + (int64_t)fq * fstride;
        const cx<double> *V = vals + (int64_t)fq * vstride;
        const int f0 = root_first[r], f1 = root_last[r];
        for (int f = f0; f <= f1; f++) {
            const FrontDev F = fronts[f];
            const int m = F.m, nc = F.nchild, mm = m * m, slot = slot_of[f];
            if (tid < nc) {
                const FrontDev C = fronts[children[F.child_off + tid]];
                cbase[tid] = C.buf_off + C.p + (int64_t)C.p * C.m;   // contribution block origin
                cstride[tid] = C.m;
            }
            for (int c = 0; c < nc; c++) {
                const int io = fronts[children[F.child_off + c]].inv_off;
                for (int i = tid; i < m; i += blockDim.x) cmap[c][i] = inv[io + i];
            }
            __syncthreads();
#pragma unroll 4
            for (int e = tid; e < mm; e += blockDim.x) {
                const int c2 = e / m, i = e - c2 * m;
                cx<T> v = {T(0), T(0)};
                for (int c = 0; c < nc; c++) {
                    const int a = cmap[c][i], b = cmap[c][c2];
                    if (a >= 0 && b >= 0) v = cadd(v, A[cbase[c] + a + (int64_t)b * cstride[c]]);
                }
                S[e] = v;
            }
            __syncthreads();
            for (int64_t e = optr[slot] + tid; e < optr[slot + 1]; e += blockDim.x) {
                const cx<double> v = V[oq[e]];
                cx<T> &d = S[ooff[e]];
                d.x += (T)v.x;
                d.y += (T)v.y;
            }
            __syncthreads();
            small_lu_lds<T>(S, m, F.p, piv + (int64_t)fq * n + F.pos0, sing + fq, sp);
            cx<T> *G = A + F.buf_off;
            for (int e = tid; e < mm; e += blockDim.x) G[e] = S[e];
            __syncthreads();   // the parent, later in this loop, reads this block
        }
    }
}

// ---------------------------------------------------------------------------
// 3b: blocked path for large fronts
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_panel(cx<T> *__restrict__ arena, int64_t fstride,
                                               const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                               int *__restrict__ piv, int n, int k0, int nb, int *__restrict__ sing) {
    __shared__ T sb[4];
    __shared__ int si[4];
    const FrontDev F = fronts[list[blockIdx.x]];
    if (F.p <= k0) return;
    const int fq = blockIdx.y;
    cx<T> *G = arena + fq * fstride + F.buf_off;
    const int m = F.m, p = F.p, kb = min(nb, p - k0), tid = threadIdx.x, bs = blockDim.x;
    for (int j = k0; j < k0 + kb; j++) {
        T best = T(-1);
        int bi = j;
        for (int r = j + tid; r < p; r += bs) better(best, bi, cabs1(G[r + (int64_t)j * m]), r);
        const int pr = block_argmax(best, bi, sb, si);
        if (tid == 0) {
            piv[(int64_t)fq * n + F.pos0 + j] = pr;
            if (sb[0] == T(0)) atomicOr(sing + fq, 1);
        }
        if (pr != j)
            for (int c = tid; c < m; c += bs) {
                cx<T> t = G[j + (int64_t)c * m];
                G[j + (int64_t)c * m] = G[pr + (int64_t)c * m];
                G[pr + (int64_t)c * m] = t;
            }
        __syncthreads();
        const cx<T> d = crecip(G[j + (int64_t)j * m]);
        for (int r = j + 1 + tid; r < m; r += bs) G[r + (int64_t)j * m] = cmul(G[r + (int64_t)j * m], d);
        __syncthreads();
        const int rows = m - j - 1, ncol = k0 + kb - j - 1;
        for (int e = tid; e < rows * ncol; e += bs) {
            int c = j + 1 + e / rows, r = j + 1 + e % rows;
            G[r + (int64_t)c * m] = cmsub(G[r + (int64_t)c * m], G[r + (int64_t)j * m], G[j + (int64_t)c * m]);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 3c: blocked path, panel held in LDS.  The pivot candidates of a 32-column
// panel are the front's remaining fully-summed rows [k0, p); their panel
// columns are factored entirely in LDS (column-major, np rows), with the row
// swaps kept inside the panel.  k_panel_apply then applies the swap sequence to
// the front's other columns and finishes the panel columns of the
// contribution rows [p, m), which need no pivoting: row a of them becomes the
// solution of x U_kk = a.  Same result as k_panel, with the per-column steps in
// LDS instead of global memory.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_panel_lu(cx<T> *__restrict__ arena, int64_t fstride,
                                                  const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                  int *__restrict__ piv, int n, int k0, int nb, int *__restrict__ sing) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *P = reinterpret_cast<cx<T> *>(smem_raw);
    __shared__ int si[1];
    const FrontDev F = fronts[list[blockIdx.x]];
    if (F.p <= k0) return;
    const int fq = blockIdx.y;
    cx<T> *G = arena + fq * fstride + F.buf_off;
    const int m = F.m, p = F.p, kb = min(nb, p - k0), np = p - k0, tid = threadIdx.x;
    // the panel staged eight entries per thread at a time, all eight loads in
    // flight before the stores (entries past the panel read its first)
    for (int e0 = 0; e0 < np * kb; e0 += 8 * 256) {
        cx<T> v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = e0 + tid + 256 * u, ec = e < np * kb ? e : 0, r = ec % np, c = ec / np;
            v[u] = G[(k0 + r) + (int64_t)(k0 + c) * m];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = e0 + tid + 256 * u;
            if (e < np * kb) P[(e / np) * np + e % np] = v[u];
        }
    }
    __syncthreads();
    __shared__ cx<T> pv[2];   // pivot entry and the entry it swaps with, published with the pivot row
    for (int j = 0; j < kb; j++) {
        // pivot search by wave 0 alone (a few candidates per lane)
        if (tid < 64) {
            T best = T(-1);
            int bi = j;
            for (int r = j + tid; r < np; r += 64) better(best, bi, cabs1(P[j * np + r]), r);
            for (int off = 32; off > 0; off >>= 1) {
                const T ob = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                better(best, bi, ob, oi);
            }
            if (tid == 0) {
                si[0] = bi;
                pv[0] = P[j * np + bi];
                pv[1] = P[j * np + j];
                piv[(int64_t)fq * n + F.pos0 + k0 + j] = k0 + bi;
                if (best == T(0)) atomicOr(sing + fq, 1);
            }
        }
        __syncthreads();
        // row swap in the panel's other columns and the scaled column j, in one
        // pass: column j's swapped pair comes from pv, read before anyone wrote
        const int pr = si[0];
        const cx<T> d = crecip(pv[0]);
        if (pr != j && tid < kb && tid != j) {
            const cx<T> t = P[tid * np + j];
            P[tid * np + j] = P[tid * np + pr];
            P[tid * np + pr] = t;
        }
        for (int r = j + 1 + tid; r < np; r += 256) P[j * np + r] = cmul(r == pr ? pv[1] : P[j * np + r], d);
        if (tid == 0) P[j * np + j] = pv[0];
        __syncthreads();
        const int rows = np - j - 1, ncol = kb - j - 1;
        for (int e = tid; e < rows * ncol; e += 256) {
            const int c = j + 1 + e / rows, r = j + 1 + e % rows;
            P[c * np + r] = cmsub(P[c * np + r], P[j * np + r], P[c * np + j]);
        }
        __syncthreads();
    }
    for (int e = tid; e < np * kb; e += 256) {
        const int r = e % np, c = e / np;
        G[(k0 + r) + (int64_t)(k0 + c) * m] = P[c * np + r];
    }
}

// grid (column or row chunks of 64, fronts, nfreq): chunk x < ccols handles 64
// columns outside the panel (swap sequence), the rest 64 contribution rows
template <typename T>
__global__ void __launch_bounds__(64) k_panel_apply(cx<T> *__restrict__ arena, int64_t fstride,
                                                    const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                    const int *__restrict__ piv, int n, int k0, int nb, int ccols) {
    __shared__ cx<T> U[32][33];
    __shared__ int ps[32];
    const FrontDev F = fronts[list[blockIdx.y]];
    if (F.p <= k0) return;
    const int fq = blockIdx.z;
    cx<T> *G = arena + fq * fstride + F.buf_off;
    const int m = F.m, p = F.p, kb = min(nb, p - k0), tid = threadIdx.x;
    if ((int)blockIdx.x < ccols) {
        // columns c outside [k0, k0+kb): the panel's row swaps, in order
        if (tid < kb) ps[tid] = piv[(int64_t)fq * n + F.pos0 + k0 + tid];
        __syncthreads();
        int c = blockIdx.x * 64 + tid;
        if (c >= k0) c += kb;
        if (c >= m) return;
        cx<T> *col = G + (int64_t)c * m;
        for (int j = 0; j < kb; j++) {
            const int r = ps[j];
            if (r != k0 + j) {
                const cx<T> t = col[k0 + j];
                col[k0 + j] = col[r];
                col[r] = t;
            }
        }
        return;
    }
    // contribution rows: x U_kk = a, U_kk from the factored panel
    // the block U_kk (16 entries per thread, all in flight before the stores),
    // then the row's panel entries (unconditional: columns past kb read column
    // k0 and are not used)
    const int i = p + (blockIdx.x - ccols) * 64 + tid, ic = i < m ? i : p;
    {
        cx<T> uv[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = tid + 64 * u, r = e & 31, c = e >> 5;
            uv[u] = G[(k0 + (r < kb ? r : 0)) + (int64_t)(k0 + (c < kb ? c : 0)) * m];
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = tid + 64 * u, r = e & 31, c = e >> 5;
            if (r < kb && c < kb) U[r][c] = uv[u];
        }
    }
    __syncthreads();
    if (i >= m) return;
    cx<T> x[32];
#pragma unroll
    for (int t = 0; t < 32; t++) x[t] = G[ic + (int64_t)(k0 + (t < kb ? t : 0)) * m];
#pragma unroll
    for (int t = 0; t < 32; t++) {
        if (t < kb) {
            cx<T> a = x[t];
#pragma unroll
            for (int s = 0; s < t; s++) a = cmsub(a, x[s], U[s][t]);
            x[t] = cmul(a, crecip(U[t][t]));
        }
    }
#pragma unroll
    for (int t = 0; t < 32; t++)
        if (t < kb) G[i + (int64_t)(k0 + t) * m] = x[t];
}

// U12 rows of the pivot block: rows [k0, k0+kb) of the columns beyond it,
// solved with the unit lower triangle L[k0:k0+kb, k0:k0+kb].  KB = 32: one
// panel, kb = min(32, p - k0), XW = 64 columns per workgroup; KB = 64: two
// panels at once, kb = min(64, p - k0), 32 columns per workgroup, triangle and
// columns in dynamic LDS.  ccap > 0 ends the columns at min(k0 + ccap, p) (the
// second panel's columns only), else at m.  grid (column chunks, fronts, nfreq)
template <typename T, int KB>
__global__ void __launch_bounds__(256) k_trsm_u(cx<T> *__restrict__ arena, int64_t fstride,
                                               const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                               int k0, int ccap) {
    constexpr int XW = KB == 32 ? 64 : 32, LPC = 256 / XW;   // columns per workgroup, lanes per column
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>(*Ls)[KB + 1] = reinterpret_cast<cx<T>(*)[KB + 1]>(smem_raw);
    cx<T>(*Xs)[XW + 1] = reinterpret_cast<cx<T>(*)[XW + 1]>(smem_raw + sizeof(cx<T>) * KB * (KB + 1));
    const FrontDev F = fronts[list[blockIdx.y]];
    if (F.p <= k0) return;
    const int m = F.m, kb = min(KB, F.p - k0), cend = ccap > 0 ? min(k0 + ccap, F.p) : m;
    const int c0 = k0 + kb + blockIdx.x * XW;
    if (c0 >= cend) return;
    cx<T> *G = arena + blockIdx.z * fstride + F.buf_off;
    const int tid = threadIdx.x;
    // unrolled so that a batch of loads is in flight at once
#pragma unroll 4
    for (int e = tid; e < kb * kb; e += 256) {
        int i = e % kb, t = e / kb;
        Ls[i][t] = G[(k0 + i) + (int64_t)(k0 + t) * m];
    }
    // (unconditional: columns past the end read column c0 and are zeroed when stored)
    {
        constexpr int PER = KB * XW / 256;
        cx<T> v[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = tid + 256 * u, i = e % KB, c = e / KB, col = c0 + c;
            v[u] = G[(k0 + (i < kb ? i : 0)) + (int64_t)(col < cend ? col : c0) * m];
        }
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = tid + 256 * u, i = e % KB, c = e / KB, col = c0 + c;
            if (i < kb) Xs[i][c] = col < cend ? v[u] : cx<T>{T(0), T(0)};
        }
    }
    __syncthreads();
    {
        // LPC adjacent lanes (one wave) per column: each sums every LPC-th term
        // of row i's dot product, xor-shuffles add them, lane 0 of the group
        // stores the entry the next rows read
        const int c = tid / LPC, part = tid % LPC;
        for (int i = 1; i < kb; i++) {
            cx<T> acc = {T(0), T(0)};
            for (int t = part; t < i; t += LPC) acc = cmsub(acc, Ls[i][t], Xs[t][c]);
#pragma unroll
            for (int o = 1; o < LPC; o <<= 1) {
                acc.x += __shfl_xor(acc.x, o);
                acc.y += __shfl_xor(acc.y, o);
            }
            if (part == 0) Xs[i][c] = cadd(Xs[i][c], acc);
        }
    }
    __syncthreads();
    for (int e = tid; e < kb * XW; e += 256) {
        int i = e % kb, c = e / kb, col = c0 + c;
        if (col < cend) G[(k0 + i) + (int64_t)col * m] = Xs[i][c];
    }
}
template <typename T, int KB>
constexpr size_t trsm_lds() {
    return sizeof(cx<T>) * KB * (KB + 1) + sizeof(cx<T>) * KB * ((KB == 32 ? 64 : 32) + 1);
}

// ---------------------------------------------------------------------------
// trailing update C -= A*B on MFMA (complex = 3 real products in FP32, 4 in FP64), 64x64 tiles.
// The product's depth is the front's next min(depth, p - k0) pivot columns
// (depth 32: one panel; 64: two panels), taken 32 at a time through the same
// LDS while the C tile stays in the accumulators, so C is read and written once
// per `depth` pivots.  The updated block starts at row and column k0 + that
// depth; ccap > 0 ends its columns at min(k0 + ccap, p) (the thin update of the
// second panel's columns), else at m.
// DEEP = false compiles the one-chunk form (depth <= 32) without the second
// chunk's prefetch registers, which would cost occupancy.
// 4 waves in 2x2, each wave 32x32 = 2x2 blocks of 16x16.
// MFMA 16x16x4 operand layout: A lane l -> (row l%16, k l/16), B lane l ->
// (k l/16, col l%16).  Accumulator register v of lane l holds column l%16 and
// row 4*(l/16)+v for f32, but row 4*v+l/16 for f64.
// ---------------------------------------------------------------------------
template <typename T, bool DEEP>
__global__ void __launch_bounds__(256) k_gemm_update(cx<T> *__restrict__ arena, int64_t fstride,
                                                     const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                     const TileDev *__restrict__ tiles, int k0, int depth, int ccap) {
    constexpr int KB = 32;
    __shared__ T Ar[KB][64], Ai[KB][64], Br[KB][64], Bi[KB][64];
    const TileDev tl = tiles[blockIdx.x];
    const FrontDev F = fronts[list[tl.front]];
    const int m = F.m, kd = min(depth, F.p - k0);
    const int base = k0 + kd, cend = ccap > 0 ? min(k0 + ccap, F.p) : m;
    const int r0 = base + tl.ti * 64, c0 = base + tl.tj * 64;
    cx<T> *G = arena + blockIdx.y * fstride + F.buf_off;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6, wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    // the lane's 16 entries of C: loaded up front with the first operand chunk
    // (one chunk), or after the products (two chunks: registers for occupancy)
    cx<T> c[2][2][4];
    auto load_c = [&]() {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    // FP64: unconditional (entries outside the block read G[0] and are never
                    // stored); FP32 keeps the conditional form, which measured as fast
                    const int row = r0 + wr + a * 16 + mfma_t<T>::row(lk, v), col = c0 + wc + b * 16 + lr;
                    const bool in = row < m && col < cend;
                    if constexpr (sizeof(T) == 8) c[a][b][v] = G[in ? row + (int64_t)col * m : 0];
                    else c[a][b][v] = in ? G[row + (int64_t)col * m] : cx<T>{T(0), T(0)};
                }
    };
    if (!DEEP) load_c();
    cx<T> av[KB * 64 / 256], bv[KB * 64 / 256];
    // operand columns/rows k0+kc .. k0+kc+31 into registers; FP64 loads them
    // unconditionally (entries outside read G[0]) so they stay in flight
    // (19.4 against 20.8 ms per factorisation), FP32 conditionally (measured as
    // fast); stage_chunk zeroes the entries outside as it stores them
    int lkb = 0;   // columns of the chunk in av/bv
    auto load_chunk = [&](int kc) {
        const int kb = min(KB, kd - kc);
        lkb = kb;
#pragma unroll
        for (int u = 0; u < KB * 64 / 256; u++) {
            const int e = tid + 256 * u, i = e & 63, k = e >> 6;
            const bool in = k < kb && r0 + i < m;
            if constexpr (sizeof(T) == 8) av[u] = G[in ? (r0 + i) + (int64_t)(k0 + kc + k) * m : 0];
            else av[u] = in ? G[(r0 + i) + (int64_t)(k0 + kc + k) * m] : cx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int u = 0; u < KB * 64 / 256; u++) {
            const int e = tid + 256 * u, k = e % KB, j = e / KB;
            const bool in = k < kb && c0 + j < cend;
            if constexpr (sizeof(T) == 8) bv[u] = G[in ? (k0 + kc + k) + (int64_t)(c0 + j) * m : 0];
            else bv[u] = in ? G[(k0 + kc + k) + (int64_t)(c0 + j) * m] : cx<T>{T(0), T(0)};
        }
    };
    auto stage_chunk = [&]() {
#pragma unroll
        for (int u = 0; u < KB * 64 / 256; u++) {
            const int e = tid + 256 * u, ia = e & 63, ka = e >> 6, kbb = e % KB, jb = e / KB;
            const bool ina = ka < lkb && r0 + ia < m, inb = kbb < lkb && c0 + jb < cend;
            Ar[ka][ia] = ina ? av[u].x : T(0);
            Ai[ka][ia] = ina ? av[u].y : T(0);
            Br[kbb][jb] = inb ? bv[u].x : T(0);
            Bi[kbb][jb] = inb ? bv[u].y : T(0);
        }
    };
    cacc3<T> acc[2][2];   // three real products per complex product
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b].zero();
    load_chunk(0);
    stage_chunk();
    __syncthreads();
    // FP32: the next chunk's loads in flight during this chunk's products; FP64
    // loads it after them (its prefetch registers would halve the occupancy)
    constexpr bool kPrefetch = sizeof(T) == 4;
    for (int kc = 0; kc < kd; kc += KB) {
        const bool more = DEEP && kc + KB < kd;
        if (kPrefetch && more) load_chunk(kc + KB);
        const int kend = (min(KB, kd - kc) + 3) & ~3;
        for (int kk = 0; kk < kend; kk += 4) {
            const int k = kk + lk;
            T ar[2], ai[2], br[2], bi[2];
            for (int a = 0; a < 2; a++) {
                ar[a] = Ar[k][wr + a * 16 + lr];
                ai[a] = Ai[k][wr + a * 16 + lr];
            }
            for (int b = 0; b < 2; b++) {
                br[b] = Br[k][wc + b * 16 + lr];
                bi[b] = Bi[k][wc + b * 16 + lr];
            }
            for (int b = 0; b < 2; b++) {
                const T bs = br[b] + bi[b];
                for (int a = 0; a < 2; a++) acc[a][b].mma(ar[a], ai[a], br[b], bi[b], bs);
            }
        }
        if (more) {
            __syncthreads();   // every wave is done with this chunk
            if (!kPrefetch) load_chunk(kc + KB);
            stage_chunk();
            __syncthreads();
        }
    }
    if (DEEP) load_c();
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = r0 + wr + a * 16 + mfma_t<T>::row(lk, v), col = c0 + wc + b * 16 + lr;
                if (row < m && col < cend)
                    G[row + (int64_t)col * m] = {c[a][b][v].x - acc[a][b].re(v), c[a][b][v].y - acc[a][b].im(v)};
            }
}

// ---------------------------------------------------------------------------
// row permutation of every front from its swap sequence: rowperm[pos0+i] is
// the local row that ends at position i
// ---------------------------------------------------------------------------
// zero-pivot flags cleared by a kernel rather than a memset node, so the
// factorisation can be captured into a graph and replayed
__global__ void k_clear_flags(int *__restrict__ flags, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) flags[i] = 0;
}

__global__ void k_rowperm(const FrontDev *__restrict__ fronts, int nfronts, const int *__restrict__ piv, int n,
                          int *__restrict__ rowperm) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfronts) return;
    const FrontDev F = fronts[f];
    const int64_t o = (int64_t)blockIdx.y * n;
    int *rp = rowperm + o + F.pos0;
    const int *pv = piv + o + F.pos0;
    for (int i = 0; i < F.p; i++) rp[i] = i;
    for (int j = 0; j < F.p; j++) {
        int t = pv[j];
        if (t != j) { int a = rp[j]; rp[j] = rp[t]; rp[t] = a; }
    }
}

// ---------------------------------------------------------------------------
// gather sources after pivoting, per frequency: row i of a front takes the
// sources of its row rowperm(i) (pivot rows) or i (contribution rows), so a
// forward sweep reads one table entry per row.  grid (row chunks, nfreq)
// ---------------------------------------------------------------------------
__global__ void k_gsrc(const FrontDev *__restrict__ fronts, int nfronts, const int *__restrict__ rowperm, int n,
                       const int *__restrict__ gsrc_id, int64_t fv_rows, int *__restrict__ gsrc) {
    const int f = blockIdx.x;
    if (f >= nfronts) return;
    const FrontDev F = fronts[f];
    const int *rp = rowperm + (int64_t)blockIdx.y * n + F.pos0;
    int *out = gsrc + ((int64_t)blockIdx.y * fv_rows + F.fv_off) * 3;
    const int *in = gsrc_id + F.fv_off * 3;
    for (int i = threadIdx.x; i < F.m; i += blockDim.x) {
        const int s = i < F.p ? rp[i] : i;
        out[3 * i] = in[3 * s];
        out[3 * i + 1] = in[3 * s + 1];
        out[3 * i + 2] = in[3 * s + 2];
    }
}

// ---------------------------------------------------------------------------
// solve format: every 32x32 diagonal block of a front's pivot block is
// replaced in place by its inverses -- strictly lower part: inverse of the
// unit lower block of L11, upper part with diagonal: inverse of the upper
// block of U11 -- so that the substitutions apply each diagonal block as one
// small dense product on the matrix cores instead of a 32-step dependent
// chain.  Fronts of single-block levels (p <= 32) go one step further: L21 is
// replaced by L21 inv(L11) and U12 by inv(U11) U12, which turns each of their
// forward and backward substitutions -- for M as for M^T -- into one product
// with no dependent phase:  [y; w2 - L21 y] = [inv(L11); L21 inv(L11)] w1  and
// x1 = [inv(U11), -inv(U11) U12] [y1; x2].  Computed in FP64 for both factor
// precisions (the block is staged in the factor precision, exact in FP64).
// One workgroup per front, grid (fronts, nfreq); every global load of the
// front is issued a phase ahead of its use (the next block, or the first
// row and column chunk of the rectangles, while a block is inverted).
// ---------------------------------------------------------------------------
// every lane takes the value that lane q of its group of four holds (DPP quad permute)
__device__ __forceinline__ int quad_bcast_i(int v, int q) {
    switch (q) {
    case 0: return __builtin_amdgcn_mov_dpp(v, 0x00, 0xF, 0xF, false);
    case 1: return __builtin_amdgcn_mov_dpp(v, 0x55, 0xF, 0xF, false);
    case 2: return __builtin_amdgcn_mov_dpp(v, 0xAA, 0xF, 0xF, false);
    default: return __builtin_amdgcn_mov_dpp(v, 0xFF, 0xF, 0xF, false);
    }
}
__device__ __forceinline__ double quad_bcast(double d, int q) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(d);
    const unsigned lo = (unsigned)quad_bcast_i((int)(unsigned)b, q);
    const unsigned hi = (unsigned)quad_bcast_i((int)(unsigned)(b >> 32), q);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ cx<double> quad_bcast(cx<double> v, int q) {
    return {quad_bcast(v.x, q), quad_bcast(v.y, q)};
}
__device__ __forceinline__ double readlane_d(double d, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(d);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T>
__device__ __forceinline__ cx<double> c64(cx<T> v) {
    return {(double)v.x, (double)v.y};
}
// c + a b with every fused multiply-add written out
__device__ __forceinline__ cx<double> cmac_f(cx<double> c, cx<double> a, cx<double> b) {
    return {fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))};
}

// Inverses of the NB x NB block D (NB = 8, 16, 24 or 32, the identity beyond the
// block's order): the unit lower inverse into LI and the upper inverse into UI,
// each complete with its zeros (LI with its unit diagonal).  Waves 0-1 build
// LI, waves 2-3 UI at the same time; four lanes per column j, lane part holding
// rows 4k + part in registers.  Right-looking: at step t the lane holding row
// t hands x_t to its group of four (DPP), and every lane updates its rows below
// t (above t for UI) -- no LDS round trip and no barrier inside the chain.
template <typename T, int NB>
__device__ __forceinline__ void tri_inverses(const cx<T> (*D)[33], cx<double> (*LI)[33], cx<double> (*UI)[33],
                                             int tid) {
    constexpr int NK = NB / 4;
    const int j = (tid & 127) >> 2, part = tid & 3;
    cx<double> acc[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) acc[k] = {4 * k + part == j ? 1.0 : 0.0, 0.0};
    // column t of the block for the lane's rows, read a step ahead; the
    // scheduling barrier per step keeps the unrolled chain from hoisting every
    // LDS read at once
    cx<T> an[NK];
    if (tid < 128) {
        // column j of inv(L): x_t final once rows < t are applied; rows i > t -= L[i][t] x_t
#pragma unroll
        for (int k = 0; k < NK; k++) an[k] = D[4 * k + part][0];
#pragma unroll
        for (int t = 0; t < NB - 1; t++) {
            const int q = t & 3, kt = t >> 2;
            cx<T> ac[NK];
#pragma unroll
            for (int k = kt; k < NK; k++) ac[k] = an[k];
            if (t + 1 < NB - 1)
#pragma unroll
                for (int k = (t + 1) >> 2; k < NK; k++) an[k] = D[4 * k + part][t + 1];
            const cx<double> x = quad_bcast(acc[kt], q);
#pragma unroll
            for (int k = kt; k < NK; k++) {
                cx<double> a = c64(ac[k]);
                if (k == kt && part <= q) a = {0.0, 0.0};
                acc[k] = cmsub_f(acc[k], a, x);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (j < NB)
#pragma unroll
            for (int k = 0; k < NK; k++) LI[4 * k + part][j] = acc[k];
    } else {
        // column j of inv(U): x_t = acc_t / U[t][t] once rows > t are applied; rows i < t -= U[i][t] x_t
        // lane l holds 1 / U[l % 32][l % 32]; step t reads lane t's (readlane)
        const int dl = tid & 31;
        const cx<double> rl = crecip(c64(D[dl][dl]));
#pragma unroll
        for (int k = 0; k < NK; k++) an[k] = D[4 * k + part][NB - 1];
#pragma unroll
        for (int t = NB - 1; t >= 0; t--) {
            const int q = t & 3, kt = t >> 2;
            cx<T> ac[NK];
#pragma unroll
            for (int k = 0; k <= kt; k++) ac[k] = an[k];
            if (t > 0)
#pragma unroll
                for (int k = 0; k <= (t - 1) >> 2; k++) an[k] = D[4 * k + part][t - 1];
            const cx<double> x = quad_bcast(cmul(acc[kt], cx<double>{readlane_d(rl.x, t), readlane_d(rl.y, t)}), q);
            if (part == q) acc[kt] = x;
#pragma unroll
            for (int k = 0; k <= kt; k++) {
                cx<double> a = c64(ac[k]);
                if (k == kt && part >= q) a = {0.0, 0.0};
                acc[k] = cmsub_f(acc[k], a, x);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (j < NB)
#pragma unroll
            for (int k = 0; k < NK; k++) UI[4 * k + part][j] = acc[k];
    }
}

// L21 <- L21 inv(L11) on a chunk of 32 rows staged as V[i][s] (zero beyond the
// front), on the matrix cores: wave w computes one 16 x 16 tile of the
// transposed product (rows t: columns of L21, columns i: its rows), so each
// accumulator register stores 16 consecutive rows of one column of G.  K runs
// from the tile's first column (inv(L11) is zero above its diagonal).
template <typename T>
__device__ __forceinline__ void l21_mfma(const cx<T> (*V)[33], const cx<double> (*LI)[33], cx<T> *G, int m, int p,
                                         int r0, int wave, int lane) {
    const int ntt = (p + 15) >> 4, lr = lane & 15, lk = lane >> 4;
    if (wave >= 2 * ntt) return;
    const int t0 = (wave % ntt) * 16, i0 = (wave / ntt) * 16;
    if (r0 + i0 >= m) return;
    d4v cr = {0.0, 0.0, 0.0, 0.0}, ci = cr;
    const int kend = (p + 3) & ~3;
    for (int k0 = t0; k0 < kend; k0 += 4) {
        const cx<double> a = LI[k0 + lk][t0 + lr];       // A[t][s] = inv(L11)[s][t]
        const cx<double> b = c64(V[i0 + lr][k0 + lk]);   // B[s][i] = L21[i][s]
        cmma<double>(a.x, a.y, b.x, b.y, cr, ci);
    }
    const int i = i0 + lr;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int t = t0 + mfma_t<double>::row(lk, v);
        if (t < p && r0 + i < m) G[r0 + i + (int64_t)t * m] = {(T)cr[v], (T)ci[v]};
    }
}

// U12 <- inv(U11) U12 on a chunk of 32 columns staged as V[c][s]: wave w
// computes one 16 x 16 tile of the transposed product (rows c: columns of U12,
// columns i: its rows); K runs from the tile's first row (inv(U11) is zero
// below its diagonal).
template <typename T>
__device__ __forceinline__ void u12_mfma(const cx<T> (*V)[33], const cx<double> (*UI)[33], cx<T> *G, int m, int p,
                                         int c0, int wave, int lane) {
    const int nit = (p + 15) >> 4, lr = lane & 15, lk = lane >> 4;
    if (wave >= 2 * nit) return;
    const int i0 = (wave % nit) * 16, cc0 = (wave / nit) * 16;
    if (c0 + cc0 >= m) return;
    d4v cr = {0.0, 0.0, 0.0, 0.0}, ci = cr;
    const int kend = (p + 3) & ~3;
    for (int k0 = i0; k0 < kend; k0 += 4) {
        const cx<double> a = c64(V[cc0 + lr][k0 + lk]);   // A[c][s] = U12[s][c]
        const cx<double> b = UI[i0 + lr][k0 + lk];        // B[s][i] = inv(U11)[i][s]
        cmma<double>(a.x, a.y, b.x, b.y, cr, ci);
    }
    const int i = i0 + lr;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int c = c0 + cc0 + mfma_t<double>::row(lk, v);
        if (i < p && c < m) G[i + (int64_t)c * m] = {(T)cr[v], (T)ci[v]};
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_diag_inv(cx<T> *__restrict__ arena, int64_t fstride,
                                                  const FrontDev *__restrict__ fronts,
                                                  const int *__restrict__ lv_single) {
    __shared__ cx<T> D[32][33];   // the block being inverted, then the rectangles' chunks
    __shared__ cx<double> LI[32][33], UI[32][33];
    const FrontDev F = fronts[blockIdx.x];
    cx<T> *G = arena + blockIdx.y * fstride + F.buf_off;
    const int m = F.m, p = F.p, tid = threadIdx.x;
    if (p == 0) return;
    const bool single = lv_single[F.level] && p <= 32;
    // chunk loads: four entries per thread, entry e = tid + 256u at (e % 32,
    // e / 32) of a 32 x 32 tile; out-of-range entries read G[0] and are
    // replaced when staged (unconditional loads stay in flight together)
    cx<T> rb[4], rl[4], ru[4];
    auto load_block = [&](int b0) {
        const int kb = min(32, p - b0);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = tid + 256 * u, i = e & 31, t = e >> 5;
            rb[u] = G[i < kb && t < kb ? b0 + i + (int64_t)(b0 + t) * m : 0];
        }
    };
    auto load_l21 = [&](int r0) {   // rows r0 + i, columns s < p
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = tid + 256 * u, i = e & 31, s = e >> 5;
            rl[u] = G[r0 + i < m && s < p ? r0 + i + (int64_t)s * m : 0];
        }
    };
    auto load_u12 = [&](int c0) {   // rows s < p, columns c0 + c
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = tid + 256 * u, s = e & 31, c = e >> 5;
            ru[u] = G[s < p && c0 + c < m ? s + (int64_t)(c0 + c) * m : 0];
        }
    };
    load_block(0);
    if (single) {
        if (p < m) {
            load_l21(p);
            load_u12(p);
        }
    }
    for (int b0 = 0; b0 < p; b0 += 32) {
        const int kb = min(32, p - b0);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = tid + 256 * u, i = e & 31, t = e >> 5;
            D[i][t] = i < kb && t < kb ? rb[u] : cx<T>{i == t ? T(1) : T(0), T(0)};
        }
        __syncthreads();
        if (b0 + 32 < p) load_block(b0 + 32);
        // the block's order rounded up to 8 (the padding is the identity)
        if (kb <= 8) tri_inverses<T, 8>(D, LI, UI, tid);
        else if (kb <= 16) tri_inverses<T, 16>(D, LI, UI, tid);
        else if (kb <= 24) tri_inverses<T, 24>(D, LI, UI, tid);
        else tri_inverses<T, 32>(D, LI, UI, tid);
        __syncthreads();
        for (int e = tid; e < kb * kb; e += 256) {
            const int i = e % kb, t = e / kb;
            const cx<double> v = i > t ? LI[i][t] : UI[i][t];
            G[b0 + i + (int64_t)(b0 + t) * m] = {(T)v.x, (T)v.y};
        }
    }
    if (!single) return;
    const int wave = tid >> 6, lane = tid & 63;
    for (int r0 = p; r0 < m; r0 += 32) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = tid + 256 * u, i = e & 31, s = e >> 5;
            D[i][s] = r0 + i < m && s < p ? rl[u] : cx<T>{T(0), T(0)};
        }
        __syncthreads();
        if (r0 + 32 < m) load_l21(r0 + 32);
        l21_mfma<T>(D, LI, G, m, p, r0, wave, lane);
        __syncthreads();
    }
    for (int c0 = p; c0 < m; c0 += 32) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = tid + 256 * u, s = e & 31, c = e >> 5;
            D[c][s] = s < p && c0 + c < m ? ru[u] : cx<T>{T(0), T(0)};
        }
        __syncthreads();
        if (c0 + 32 < m) load_u12(c0 + 32);
        u12_mfma<T>(D, UI, G, m, p, c0, wave, lane);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// driver
// ---------------------------------------------------------------------------
namespace {
size_t g_panel_lds_cap = 0;   // 0: the hardware limit
int g_persistent = 1;         // small subtrees in one persistent launch
int g_deep_min_m = 0;         // blocked levels with a front of order >= this take the two-panel schedule
}
void set_factor_tuning(int panel_lds_cap, int persistent, int deep_min_m) {
    g_panel_lds_cap = panel_lds_cap > 0 ? (size_t)panel_lds_cap : 0;
    g_persistent = persistent != 0;
    g_deep_min_m = deep_min_m;
}

template <typename T>
void factor_numeric(const DevPlan &dp, Factors<T> &F, const cx<double> *vals, int64_t vstride, hipStream_t st) {
    const Plan &pl = *dp.host;
    const int nf = F.nfreq;
    F.vinv_ready = false;   // the single-vector path's inverse squares follow the new factors
    hipLaunchKernelGGL(k_clear_flags, dim3((nf + 63) / 64), dim3(64), 0, st, F.sing, nf);
    const size_t lds_small_max = 160 * 1024 - 256;
    const size_t panel_lds_hw = 160 * 1024 - 1024;    // panels of up to 640 c64 / 320 c128 candidate rows
    // panels above the cap take the global-memory panel kernel
    const size_t panel_lds_max = g_panel_lds_cap ? std::min(g_panel_lds_cap, panel_lds_hw) : panel_lds_hw;
    static bool attr_set = false;
    if (!attr_set) {
        // fronts of up to 90x90 complex doubles live in LDS (gfx950: 160 KB per CU)
        GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_small_lu<double>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_small_max));
        GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_small_lu<float>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_small_max));
        GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_panel_lu<double>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)panel_lds_hw));
        GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_panel_lu<float>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)panel_lds_hw));
        GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_subtree_factor<double>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024));
        GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_subtree_factor<float>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024));
        GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_trsm_u<double, 64>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)trsm_lds<double, 64>()));
        GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_trsm_u<float, 64>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)trsm_lds<float, 64>()));
        attr_set = true;
    }
    const bool sub = g_persistent && dp.nroots > 0 && F.sched;
    if (sub) {
        const int ntask = dp.nroots * nf;
        const size_t lds = (size_t)dp.sub_mmax * dp.sub_mmax * sizeof(cx<T>);
        if (lds > 64 * 1024) throw std::logic_error("subtree front exceeds its LDS");
        hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(64), 0, st, F.sched, 1);
        ProfScope ps(K_SUBTREE, st, dp.sub_flops * nf, 3.0 * dp.sub_elems * sizeof(cx<T>) * nf);
        // every workgroup takes tasks until none is left; a few per CU keep the CUs busy
        const int grid = std::min(ntask, 256 * 8);
        hipLaunchKernelGGL(k_subtree_factor<T>, dim3(grid), dim3(256), lds, st, F.arena, F.fstride, dp.fronts,
                           dp.root_first, dp.root_last, dp.slot_of, dp.children, dp.inv, dp.sub_optr, dp.sub_oq,
                           dp.sub_ooff, vals, vstride, dp.nroots, nf, F.piv, pl.n, F.sing, F.sched);
    }
    for (int l = 0; l < pl.nlevels; l++) {
        const LevelPlan &L = pl.levels[l];
        const DevPlan::LevelCounts &cnt = dp.counts[l];
        const double es = sizeof(cx<T>) * (double)nf;
        // the level path: every front of the level, or those outside the small subtrees
        const int n_asm = sub ? dp.n_asm_rest[l] : dp.n_all[l];
        const int mmax_asm = sub ? dp.mmax_asm_rest[l] : dp.mmax_all[l];
        const int *lv_asm = sub ? dp.lv_asm_rest[l] : dp.lv_all[l];
        if (n_asm > 0) {
            ProfScope ps(K_FRONT_ASSEMBLE, st, 0.0, (sub ? dp.asm_rest_elems[l] : cnt.asm_elems) * es);
            int64_t mm = (int64_t)mmax_asm * mmax_asm;
            // about 16 entries per thread (measured faster than 4 or 8)
            int chunks = (int)std::min<int64_t>(std::max<int64_t>(1, mm / (256 * 16)), 256);
            const size_t lds = sizeof(int) * (size_t)dp.max_children * mmax_asm;
            if (dp.max_children > kAsmMaxChildren || lds > 64 * 1024)
                throw std::logic_error("front assembly: too many children or too large a front for the row maps");
            hipLaunchKernelGGL(k_front_assemble<T>, dim3(n_asm, chunks, nf), dim3(256), lds, st, F.arena,
                               F.fstride, dp.fronts, lv_asm, dp.children, dp.inv);
        }
        const int64_t obeg = sub ? dp.orig_rest_range[l].first : L.orig_begin;
        const int64_t ocnt = sub ? dp.orig_rest_range[l].second - obeg : L.orig_end - L.orig_begin;
        if (ocnt > 0) {
            ProfScope ps(K_SCATTER, st, 0.0, (double)ocnt * (16.0 + 16.0 + 2 * sizeof(cx<T>)) * nf);
            int blocks = (int)std::min<int64_t>((ocnt + 255) / 256, 4096);
            hipLaunchKernelGGL(k_scatter_orig<T>, dim3(blocks, nf), dim3(256), 0, st, F.arena, F.fstride, vals,
                               vstride, (sub ? dp.orig_q_rest : dp.orig_q) + obeg,
                               (sub ? dp.orig_dst_rest : dp.orig_dst) + obeg, ocnt);
        }
        const int n_small = sub ? dp.n_small_rest[l] : dp.n_small[l];
        if (n_small > 0) {
            const int smm = sub ? dp.small_mmax_rest[l] : L.small_mmax;
            size_t lds = (size_t)smm * smm * sizeof(cx<T>);
            if (lds > lds_small_max) throw std::logic_error("small front exceeds LDS");
            ProfScope ps(K_SMALL_LU, st, (sub ? dp.small_rest_flops[l] : cnt.small_flops) * nf,
                         (sub ? dp.small_rest_elems[l] : cnt.small_elems) * es);
            hipLaunchKernelGGL(k_small_lu<T>, dim3(n_small, nf), dim3(256), lds, st, F.arena, F.fstride,
                               dp.fronts, sub ? dp.lv_small_rest[l] : dp.lv_small[l], F.piv, pl.n, F.sing);
        }
        if (dp.n_large[l] > 0) {
            const size_t lds32 = trsm_lds<T, 32>(), lds64 = trsm_lds<T, 64>();
            // one 32-column panel of every blocked front of the level: its LU
            // (in LDS, or in global memory when the candidate rows do not fit)
            // and the row swaps on the other columns with the contribution rows
            auto panel = [&](int k0, double flops) {
                const size_t plds = (size_t)(L.large_pmax - k0) * pl.nb * sizeof(cx<T>);
                if (plds <= panel_lds_max) {
                    {
                        ProfScope ps(K_PANEL, st, flops * nf, 0.0);
                        hipLaunchKernelGGL(k_panel_lu<T>, dim3(dp.n_large[l], nf), dim3(256), plds, st, F.arena,
                                           F.fstride, dp.fronts, dp.lv_large[l], F.piv, pl.n, k0, pl.nb, F.sing);
                    }
                    const int ccols = (L.large_mmax + 63) / 64;   // m - kb columns of any front fit
                    const int crows = (L.large_mmax + 63) / 64;
                    ProfScope ps(K_PANEL_APPLY, st, 0.0, 0.0);
                    hipLaunchKernelGGL(k_panel_apply<T>, dim3(ccols + crows, dp.n_large[l], nf), dim3(64), 0, st,
                                       F.arena, F.fstride, dp.fronts, dp.lv_large[l], F.piv, pl.n, k0, pl.nb, ccols);
                } else {
                    ProfScope ps(K_PANEL, st, flops * nf, 0.0);
                    hipLaunchKernelGGL(k_panel<T>, dim3(dp.n_large[l], nf), dim3(256), 0, st, F.arena, F.fstride,
                                       dp.fronts, dp.lv_large[l], F.piv, pl.n, k0, pl.nb, F.sing);
                }
            };
            auto gemm = [&](int64_t off, int cnt, int k0, int depth, int ccap, double flops, double elems) {
                if (cnt <= 0) return;
                ProfScope ps(K_GEMM, st, flops * nf, elems * es);
                if (depth > 32)
                    hipLaunchKernelGGL((k_gemm_update<T, true>), dim3(cnt, nf), dim3(256), 0, st, F.arena, F.fstride,
                                       dp.fronts, dp.lv_large[l], dp.tiles + off, k0, depth, ccap);
                else
                    hipLaunchKernelGGL((k_gemm_update<T, false>), dim3(cnt, nf), dim3(256), 0, st, F.arena,
                                       F.fstride, dp.fronts, dp.lv_large[l], dp.tiles + off, k0, depth, ccap);
            };
            if (L.large_mmax >= g_deep_min_m) {
                // two panels per block: panel 1; its U rows and update on panel 2's
                // columns only; panel 2; the U rows of both panels (one 64-row
                // solve) and one depth-64 trailing update, so the trailing block
                // is read and written once per 64 pivots.  Panel 2's row swaps
                // reach the not yet updated columns and panel 1's L columns alike,
                // which keeps the deferred update consistent.
                int bi = 0;
                for (int b0 = 0; b0 < L.large_pmax; b0 += 64, bi++) {
                    const auto &d = dp.deep_steps[l][bi];
                    panel(b0, d.panel1_flops);
                    if (d.thin_cnt > 0) {
                        {
                            ProfScope ps(K_TRSM, st, d.trsm_thin_flops * nf, 0.0);
                            hipLaunchKernelGGL((k_trsm_u<T, 32>), dim3(1, dp.n_large[l], nf), dim3(256),
                                               lds32, st, F.arena, F.fstride, dp.fronts, dp.lv_large[l],
                                               b0, 64);
                        }
                        gemm(d.thin_off, d.thin_cnt, b0, 32, 64, d.thin_flops, d.thin_elems);
                        panel(b0 + 32, d.panel2_flops);
                    }
                    {
                        ProfScope ps(K_TRSM, st, d.trsm_deep_flops * nf, 0.0);
                        const int colchunks = (L.large_mmax - b0 + 31) / 32;
                        hipLaunchKernelGGL((k_trsm_u<T, 64>), dim3(colchunks, dp.n_large[l], nf), dim3(256),
                                           lds64, st, F.arena, F.fstride, dp.fronts, dp.lv_large[l], b0,
                                           0);
                    }
                    gemm(d.deep_off, d.deep_cnt, b0, 64, 0, d.deep_flops, d.deep_elems);
                }
            } else {
                int si = 0;
                for (int k0 = 0; k0 < L.large_pmax; k0 += pl.nb, si++) {
                    panel(k0, cnt.panel_flops[si]);
                    {
                        ProfScope ps(K_TRSM, st, cnt.trsm_flops[si] * nf, 0.0);
                        const int colchunks = (L.large_mmax - k0 + 63) / 64;
                        hipLaunchKernelGGL((k_trsm_u<T, 32>), dim3(colchunks, dp.n_large[l], nf), dim3(256),
                                           lds32, st, F.arena, F.fstride, dp.fronts, dp.lv_large[l], k0,
                                           0);
                    }
                    const auto &sp = dp.steps[l][si];
                    gemm(sp.first, sp.second, k0, 32, 0, cnt.gemm_flops[si], cnt.gemm_elems[si]);
                }
            }
        }
    }
    const int nfr = (int)pl.fronts.size();
    hipLaunchKernelGGL(k_rowperm, dim3((nfr + 63) / 64, nf), dim3(64), 0, st, dp.fronts, nfr, F.piv, pl.n,
                       F.rowperm);
    hipLaunchKernelGGL(k_gsrc, dim3(nfr, nf), dim3(64), 0, st, dp.fronts, nfr, F.rowperm, pl.n, dp.gsrc_id, pl.fv_rows,
                       F.gsrc);
    {
        ProfScope ps(K_DIAG_INV, st, 0.0, 2.0 * dp.diag_elems * sizeof(cx<T>) * nf);
        hipLaunchKernelGGL(k_diag_inv<T>, dim3(nfr, nf), dim3(256), 0, st, F.arena, F.fstride, dp.fronts,
                           dp.lv_single);
    }
    GFDFD_HIP(hipGetLastError());
}

template void factor_numeric<double>(const DevPlan &, Factors<double> &, const cx<double> *, int64_t, hipStream_t);
template void factor_numeric<float>(const DevPlan &, Factors<float> &, const cx<double> *, int64_t, hipStream_t);

}  // namespace gfdfd
