// Geometric nested dissection of the 9-point grid and the multifrontal plan.
//
// A one-point-wide grid line separates a 9-point stencil graph, so the grid
// is bisected recursively across its longer side; a subgrid with at most
// leaf_area points becomes one leaf front.  Elimination positions are
// assigned in postorder (first half, second half, separator), which makes the
// pivots of every front contiguous.  The contribution rows of a front whose
// subgrid is R are exactly the grid points of the one-point ring around R:
// every such point lies on an ancestor separator or outside the grid, and
// every ring point is coupled to R after R's interior is eliminated.
#include "symbolic.hpp"

#include <algorithm>
#include <cstdio>
#include <functional>
#include <sstream>
#include <stdexcept>

namespace gfdfd {

namespace {

struct Rect { int i0, i1, j0, j1; };

struct Builder {
    int nx, ny;
    PlanOptions opt;
    Plan &pl;
    std::vector<Rect> rect;                  // subgrid of each front
    std::vector<std::vector<int>> piv;       // pivot variables of each front
    std::vector<std::vector<int>> kids;
    int next_pos = 0;

    Builder(int nx_, int ny_, const PlanOptions &o, Plan &p) : nx(nx_), ny(ny_), opt(o), pl(p) {}

    int make_front(const Rect &r, std::vector<int> &&pv, std::vector<int> &&ch) {
        int id = (int)pl.fronts.size();
        Front f;
        f.p = (int)pv.size();
        f.pos0 = next_pos;
        for (int v : pv) {
            pl.perm[next_pos] = v;
            pl.iperm[v] = next_pos;
            pl.front_of_pos[next_pos] = id;
            next_pos++;
        }
        int lvl = 0;
        for (int c : ch) { pl.fronts[c].parent = id; lvl = std::max(lvl, pl.fronts[c].level + 1); }
        f.level = lvl;
        pl.fronts.push_back(f);
        rect.push_back(r);
        piv.push_back(std::move(pv));
        kids.push_back(std::move(ch));
        return id;
    }

    int dissect(Rect r) {
        int w = r.i1 - r.i0, h = r.j1 - r.j0;
        if (w <= 0 || h <= 0) return -1;
        if ((int64_t)w * h <= opt.leaf_area) {
            std::vector<int> pv;
            for (int j = r.j0; j < r.j1; j++)
                for (int i = r.i0; i < r.i1; i++) pv.push_back(i + j * nx);
            return make_front(r, std::move(pv), {});
        }
        std::vector<int> ch, pv;
        if (w >= h) {
            int s = r.i0 + w / 2;
            int a = dissect({r.i0, s, r.j0, r.j1});
            int b = dissect({s + 1, r.i1, r.j0, r.j1});
            if (a >= 0) ch.push_back(a);
            if (b >= 0) ch.push_back(b);
            for (int j = r.j0; j < r.j1; j++) pv.push_back(s + j * nx);
        } else {
            int s = r.j0 + h / 2;
            int a = dissect({r.i0, r.i1, r.j0, s});
            int b = dissect({r.i0, r.i1, s + 1, r.j1});
            if (a >= 0) ch.push_back(a);
            if (b >= 0) ch.push_back(b);
            for (int i = r.i0; i < r.i1; i++) pv.push_back(i + s * nx);
        }
        return make_front(r, std::move(pv), std::move(ch));
    }
};

void finish_plan(Plan &pl, const std::vector<std::vector<int>> &contrib,
                 const std::vector<std::vector<int>> &kids, const PlanOptions &opt) {
    const int nf = (int)pl.fronts.size();
    pl.leaf_area = opt.leaf_area;
    pl.small_max = opt.small_max;
    pl.nb = opt.nb;
    // index lists, children, arena and work-panel offsets
    int64_t ioff = 0, boff = 0, foff = 0;
    pl.idx.clear();
    pl.children.clear();
    for (int f = 0; f < nf; f++) {
        Front &F = pl.fronts[f];
        F.m = F.p + (int)contrib[f].size();
        F.idx_off = ioff;
        for (int k = 0; k < F.p; k++) pl.idx.push_back(F.pos0 + k);
        for (int v : contrib[f]) pl.idx.push_back(v);
        ioff += F.m;
        F.buf_off = boff;
        boff += (int64_t)F.m * F.m;
        F.fv_off = foff;
        foff += F.m;
        F.child_off = (int)pl.children.size();
        F.nchild = (int)kids[f].size();
        for (int c : kids[f]) pl.children.push_back(c);
        pl.max_m = std::max(pl.max_m, F.m);
        pl.max_p = std::max(pl.max_p, F.p);
        pl.nlevels = std::max(pl.nlevels, F.level + 1);
        double fl = 0;
        for (int k = 0; k < F.p; k++) {
            double t = F.m - k - 1;
            fl += 4.0 * (2.0 * t * t + t);
        }
        pl.factor_flops += fl;
        pl.nnz_lu += (double)F.p * (2.0 * F.m - F.p);
    }
    pl.arena = boff;
    pl.fv_rows = foff;
    // child -> parent inverse maps
    int64_t inv_total = 0;
    for (int f = 0; f < nf; f++)
        if (pl.fronts[f].parent >= 0) {
            pl.fronts[f].inv_off = inv_total;
            inv_total += pl.fronts[pl.fronts[f].parent].m;
        }
    pl.inv.assign(inv_total, -1);
    for (int f = 0; f < nf; f++) {
        const Front &C = pl.fronts[f];
        if (C.parent < 0) continue;
        const Front &P = pl.fronts[C.parent];
        const int *ci = &pl.idx[C.idx_off];
        const int *pi = &pl.idx[P.idx_off];
        int a = C.p;
        for (int r = 0; r < P.m && a < C.m; r++) {
            if (pi[r] == ci[a]) { pl.inv[C.inv_off + r] = a - C.p; a++; }
            else if (pi[r] > ci[a]) throw std::logic_error("contribution row missing from parent front");
        }
        if (a != C.m) throw std::logic_error("contribution rows not contained in parent front");
    }
    // levels
    pl.levels.assign(pl.nlevels, LevelPlan());
    for (int f = 0; f < nf; f++) {
        const Front &F = pl.fronts[f];
        LevelPlan &L = pl.levels[F.level];
        if (F.m <= opt.small_max) {
            L.small.push_back(f);
            L.small_mmax = std::max(L.small_mmax, F.m);
        } else {
            L.large.push_back(f);
            L.large_pmax = std::max(L.large_pmax, F.p);
            L.large_mmax = std::max(L.large_mmax, F.m);
        }
    }
    // original entries of M, bucketed by the level of their owning front
    std::vector<std::vector<std::pair<int64_t, int64_t>>> bucket(pl.nlevels);
    for (int c = 0; c < pl.n; c++)
        for (int q = pl.colptr[c]; q < pl.colptr[c + 1]; q++) {
            int r = pl.rowind[q];
            int pr = pl.iperm[r], pc = pl.iperm[c];
            int f = pl.front_of_pos[std::min(pr, pc)];
            const Front &F = pl.fronts[f];
            const int *fi = &pl.idx[F.idx_off];
            int lr = (int)(std::lower_bound(fi, fi + F.m, pr) - fi);
            int lc = (int)(std::lower_bound(fi, fi + F.m, pc) - fi);
            if (lr >= F.m || lc >= F.m || fi[lr] != pr || fi[lc] != pc)
                throw std::logic_error("matrix entry outside its front");
            bucket[F.level].push_back({q, F.buf_off + lr + (int64_t)lc * F.m});
        }
    pl.orig_q.clear();
    pl.orig_dst.clear();
    for (int l = 0; l < pl.nlevels; l++) {
        pl.levels[l].orig_begin = (int64_t)pl.orig_q.size();
        for (auto &e : bucket[l]) { pl.orig_q.push_back(e.first); pl.orig_dst.push_back(e.second); }
        pl.levels[l].orig_end = (int64_t)pl.orig_q.size();
    }
}

}  // namespace

void grid_pattern(int nx, int ny, std::vector<int> &colptr, std::vector<int> &rowind) {
    const int n = nx * ny;
    colptr.assign(n + 1, 0);
    rowind.clear();
    rowind.reserve((size_t)9 * n);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            for (int dj = -1; dj <= 1; dj++)
                for (int di = -1; di <= 1; di++) {
                    int ii = i + di, jj = j + dj;
                    if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) rowind.push_back(ii + jj * nx);
                }
            colptr[i + j * nx + 1] = (int)rowind.size();
        }
}

int detect_grid(int n, const int *colptr, const int *rowind, int *ny_out) {
    if (n < 4) return 0;
    int c0 = colptr[1] - colptr[0];
    if (c0 != 4) return 0;
    int nx = rowind[colptr[0] + 2];
    if (nx < 2 || n % nx != 0) return 0;
    int ny = n / nx;
    if (ny < 2) return 0;
    int64_t nnz = 9LL * n - 6LL * nx - 6LL * ny + 4;
    if (colptr[n] - colptr[0] != nnz) return 0;
    std::vector<int> cp, ri;
    grid_pattern(nx, ny, cp, ri);
    for (int c = 0; c <= n; c++)
        if (colptr[c] - colptr[0] != cp[c]) return 0;
    for (int64_t q = 0; q < nnz; q++)
        if (rowind[colptr[0] + q] != ri[q]) return 0;
    *ny_out = ny;
    return nx;
}

Plan build_grid_plan(int nx, int ny, const PlanOptions &opt) {
    Plan pl;
    pl.nx = nx;
    pl.ny = ny;
    pl.n = nx * ny;
    pl.grid = true;
    pl.perm.assign(pl.n, -1);
    pl.iperm.assign(pl.n, -1);
    pl.front_of_pos.assign(pl.n, -1);
    grid_pattern(nx, ny, pl.colptr, pl.rowind);
    pl.nnz = (int64_t)pl.rowind.size();
    Builder b(nx, ny, opt, pl);
    b.dissect({0, nx, 0, ny});
    if (b.next_pos != pl.n) throw std::logic_error("dissection did not cover the grid");
    const int nf = (int)pl.fronts.size();
    std::vector<std::vector<int>> contrib(nf);
    for (int f = 0; f < nf; f++) {
        const Rect &r = b.rect[f];
        std::vector<int> &cs = contrib[f];
        for (int j = r.j0 - 1; j <= r.j1; j++) {
            if (j < 0 || j >= ny) continue;
            for (int i = r.i0 - 1; i <= r.i1; i++) {
                if (i < 0 || i >= nx) continue;
                if (i >= r.i0 && i < r.i1 && j >= r.j0 && j < r.j1) continue;
                int pos = pl.iperm[i + j * nx];
                if (pos < pl.fronts[f].pos0 + pl.fronts[f].p)
                    throw std::logic_error("ring point eliminated before its front");
                cs.push_back(pos);
            }
        }
        std::sort(cs.begin(), cs.end());
    }
    finish_plan(pl, contrib, b.kids, opt);
    return pl;
}

Plan build_dense_plan(int n, const int *colptr, const int *rowind, const PlanOptions &opt) {
    Plan pl;
    pl.nx = n;
    pl.ny = 1;
    pl.n = n;
    pl.grid = false;
    pl.perm.resize(n);
    pl.iperm.resize(n);
    pl.front_of_pos.assign(n, 0);
    for (int i = 0; i < n; i++) { pl.perm[i] = i; pl.iperm[i] = i; }
    pl.colptr.assign(colptr, colptr + n + 1);
    pl.rowind.assign(rowind + colptr[0], rowind + colptr[n]);
    for (auto &c : pl.colptr) c -= colptr[0];
    pl.nnz = (int64_t)pl.rowind.size();
    Front f;
    f.p = n;
    f.pos0 = 0;
    pl.fronts.push_back(f);
    std::vector<std::vector<int>> contrib(1), kids(1);
    PlanOptions o = opt;
    finish_plan(pl, contrib, kids, o);
    return pl;
}

void step_tiles(const Plan &pl, const LevelPlan &lv, int k0, int tile, std::vector<Tile> &out) {
    out.clear();
    for (int li = 0; li < (int)lv.large.size(); li++) {
        const Front &F = pl.fronts[lv.large[li]];
        if (F.p <= k0) continue;
        int kb = std::min(pl.nb, F.p - k0);
        int t = F.m - k0 - kb;
        if (t <= 0) continue;
        int nt = (t + tile - 1) / tile;
        for (int tj = 0; tj < nt; tj++)
            for (int ti = 0; ti < nt; ti++) out.push_back({li, ti, tj});
    }
}

std::string plan_summary(const Plan &pl) {
    std::ostringstream s;
    s << "plan " << pl.nx << "x" << pl.ny << " n=" << pl.n << " nnz=" << pl.nnz << " fronts=" << pl.fronts.size()
      << " levels=" << pl.nlevels << " max_m=" << pl.max_m << " max_p=" << pl.max_p
      << " arena=" << pl.arena << " factor_GF=" << pl.factor_flops * 1e-9 << " nnzLU=" << pl.nnz_lu;
    return s.str();
}

}  // namespace gfdfd
