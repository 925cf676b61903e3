// Symbolic analysis of the 9-point FDFD impedance matrix: geometric nested
// dissection of the structured nx*ny grid and the multifrontal plan the
// device factorisation and solves execute.
//
// Replaces the ordering/analysis half of umfpack_zi_symbolic as GERMAINE
// calls it (src/AC/forward_shot_AC.c:58).  The pattern is the same for every
// frequency and FWI iteration, so plans are cached by grid size.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gfdfd {

struct Front {
    int p = 0;           // pivots (fully summed variables)
    int m = 0;           // front order: p pivots followed by m-p contribution rows
    int pos0 = 0;        // elimination position of the first pivot (pivots are contiguous)
    int parent = -1;     // parent front, -1 for a root
    int level = 0;       // height above the leaves (leaves 0)
    int64_t idx_off = 0; // idx[idx_off .. idx_off+m): elimination positions of the rows, ascending
    int64_t buf_off = 0; // front buffer offset (elements) in the per-frequency arena, column-major m*m
    int64_t fv_off = 0;  // offset (rows) of the front's slice in the solve work panel
    int child_off = 0;   // children[child_off .. child_off+nchild)
    int nchild = 0;
    int64_t inv_off = 0; // (child fronts) inv[inv_off .. +m_parent): child-local contribution index or -1
};

// One tile of a blocked trailing update: front (index into the level's large list), tile row/col.
struct Tile { int front; int ti; int tj; };

struct LevelPlan {
    std::vector<int> small;      // fronts factored entirely in LDS by one workgroup
    std::vector<int> large;      // fronts factored by the blocked (panel/trsm/gemm) path
    int small_mmax = 0;
    int large_pmax = 0;
    int large_mmax = 0;
    int64_t orig_begin = 0, orig_end = 0;   // range in the level-sorted original-entry list
};

struct Plan {
    int nx = 0, ny = 0, n = 0;
    int64_t nnz = 0;
    bool grid = true;                // false: one dense front (fallback for small general patterns)
    std::vector<int> perm;           // position -> variable
    std::vector<int> iperm;          // variable -> position
    std::vector<Front> fronts;       // postorder (children before parents)
    std::vector<int> children;
    std::vector<int> idx;            // positions (see Front::idx_off)
    std::vector<int> inv;            // child->parent inverse maps (see Front::inv_off)
    std::vector<int> front_of_pos;   // owning front of each pivot position
    std::vector<LevelPlan> levels;
    // original entries of M (CSC order q) sorted by the level of the owning front
    std::vector<int64_t> orig_q;     // CSC index
    std::vector<int64_t> orig_dst;   // destination offset in the arena
    std::vector<int> colptr;         // canonical CSC column pointer of the 9-point pattern
    std::vector<int> rowind;         // canonical CSC row indices
    int64_t arena = 0;               // elements per frequency (sum m^2)
    int64_t fv_rows = 0;             // sum m
    int max_m = 0, max_p = 0, nlevels = 0;
    double factor_flops = 0;         // real flops of the complex partial LUs
    double nnz_lu = 0;               // entries of L and U (dense front blocks)
    int leaf_area = 0, small_max = 0, nb = 0;
};

struct PlanOptions {
    int leaf_area = 32;   // stop dissecting when a subgrid has at most this many points
    int small_max = 64;   // fronts with m <= small_max are factored in LDS
    int nb = 32;          // panel width of the blocked path
};

// Build the plan for an nx*ny 9-point grid (nx, ny >= 1).
Plan build_grid_plan(int nx, int ny, const PlanOptions &opt);

// Dense single-front plan for an arbitrary n*n pattern (colptr/rowind, 0-based, sorted).
Plan build_dense_plan(int n, const int *colptr, const int *rowind, const PlanOptions &opt);

// Canonical CSC pattern of the 9-point grid matrix (rows ascending per column).
void grid_pattern(int nx, int ny, std::vector<int> &colptr, std::vector<int> &rowind);

// Detect an nx*ny 9-point grid pattern; returns nx (and ny) or 0 when the pattern is not one.
int detect_grid(int n, const int *colptr, const int *rowind, int *ny_out);

// Tiles of the blocked trailing update for panel step k0 of a level's large fronts.
void step_tiles(const Plan &pl, const LevelPlan &lv, int k0, int tile, std::vector<Tile> &out);

std::string plan_summary(const Plan &pl);

}  // namespace gfdfd
