// Device-resident plan, factors and the host-side drivers of the numeric
// factorisation and the panel solves.
#pragma once
#include <memory>
#include <utility>
#include <vector>

#include "common.hpp"
#include "symbolic.hpp"

namespace gfdfd {

constexpr int kSingleMmax = 300;   // largest front order of a single-block level (its panels fit in LDS)
constexpr int kVecSplitFronts = 1 << 30;   // single-vector solves split the levels of fewer multi-block fronts (every level)
constexpr int kVinvPmax = 560;         // largest pivot block the inversion kernel holds in LDS (FP64)
constexpr int kTile = 64;       // output tile of the trailing-update GEMM
constexpr int kMaxBatch = 64;   // frequencies per batched solve
constexpr int kSubMaxM = 64;    // largest front order of the persistent subtree launch
constexpr int kAsmMaxChildren = 8;   // children a front may have (assembly row maps in LDS)

// Plan arrays uploaded once per grid size and shared by every factorisation.
struct DevPlan {
    std::shared_ptr<const Plan> host;
    FrontDev *fronts = nullptr;
    int *children = nullptr, *idx = nullptr, *inv = nullptr, *perm = nullptr, *iperm = nullptr;
    int64_t *orig_q = nullptr, *orig_dst = nullptr;
    int *colptr = nullptr, *rowind = nullptr;
    int64_t *tq = nullptr;                 // CSR-of-M position -> CSC index (symmetric pattern)
    std::vector<int *> lv_small, lv_large, lv_all;
    std::vector<int> n_small, n_large, n_all, mmax_all, pmax_all;
    TileDev *tiles = nullptr;
    std::vector<std::vector<std::pair<int64_t, int>>> steps;   // [level][step] -> (offset, count)
    // the two-panel schedule of the blocked path: per level and 64-column block,
    // the thin update of the second panel's columns by the first panel and the
    // depth-64 trailing update, as tile ranges of `tiles`, with their counts
    struct DeepStep {
        int64_t thin_off = 0, deep_off = 0;
        int thin_cnt = 0, deep_cnt = 0;
        double thin_flops = 0, thin_elems = 0, deep_flops = 0, deep_elems = 0;
        double panel1_flops = 0, panel2_flops = 0, trsm_thin_flops = 0, trsm_deep_flops = 0;
    };
    std::vector<std::vector<DeepStep>> deep_steps;
    size_t bytes = 0;
    // algorithmic counts per level (complex entries or real flops), for measurement
    struct LevelCounts {
        double small_flops = 0, small_elems = 0;     // partial LUs in LDS; front entries read+written
        double asm_elems = 0;                        // front entries written by the assembly
        double tri_elems = 0;                        // sum p^2/2 + p(m-p): factor entries one sweep reads
        double sum_m = 0, sum_p = 0;
        double contrib_elems = 0;                    // sum p(m-p): U12 entries of the backward contribution
        std::vector<double> gemm_flops, gemm_elems, panel_flops, trsm_flops;   // per blocked step
        std::vector<double> fwd_blk_elems, bwd_blk_elems;   // per 32-pivot block: factor entries the block reads
    };
    std::vector<LevelCounts> counts;
    double diag_elems = 0;                          // entries of all 32x32 diagonal blocks
    int max_children = 0;                           // largest number of children of one front
    std::vector<int> single;                        // per level: 1 if stored in the single-block solve format
    int *lv_single = nullptr;                       // device copy of `single`
    // single-vector solves of the multi-block levels: explicit inverses of each
    // front's pivot-block triangles (a p x p square per front: inv(L11) strictly
    // below the diagonal, inv(U11) on and above it) followed by the rectangles
    // L21 inv(L11) ((m - p) x p) and inv(U11) U12 (p x (m - p)), all
    // column-major, so each sweep of such a front is one product spread over
    // many workgroups
    std::vector<char> vinv;                         // per level: its fronts have inverses
    std::vector<int64_t *> vinv_off;                // per level (device): square offset of each front of lv_all
    int64_t vinv_elems = 0;                         // square and rectangle entries per frequency
    int *vinv_list = nullptr;                       // device: every front with a square
    int64_t *vinv_list_off = nullptr;               // device: its square's offset
    int vinv_n = 0, vinv_pmax = 0;
    int vinv_rect_tiles = 0;                        // most 64 x 64 rectangle tiles of one front
    int *gsrc_id = nullptr;                         // [fv_rows][3] gather sources of front rows before pivoting
    std::vector<FrontDev *> lv_fr;                  // per level: the level's fronts, contiguous
    std::vector<int> lv_children;                   // per level: 1 if any front has children

    // Small subtrees (fronts whose whole subtree has m <= small_max), factored
    // by one persistent launch: a task is one maximal such subtree (a
    // contiguous postorder range of fronts) for one frequency, taken from a
    // global counter by a workgroup that factors its fronts leaves to root.
    // The level path takes the remaining fronts.
    int nsub = 0, nroots = 0, sub_mmax = 0;
    int *root_first = nullptr, *root_last = nullptr;   // [nroots] front range of each subtree
    int *slot_of = nullptr;                         // [nfronts] subtree-front slot, -1 outside the subtrees
    int64_t *sub_optr = nullptr;                    // [nsub + 1] ranges of the slots' original entries
    int64_t *sub_oq = nullptr;                      // CSC index of each such entry
    int *sub_ooff = nullptr;                        // its local offset r + c m in the front
    double sub_flops = 0, sub_elems = 0;            // algorithmic counts of the launch (per frequency)
    std::vector<int *> lv_asm_rest, lv_small_rest;  // per level: fronts outside the subtrees
    std::vector<int> n_asm_rest, n_small_rest, mmax_asm_rest, small_mmax_rest;
    int64_t *orig_q_rest = nullptr, *orig_dst_rest = nullptr;
    std::vector<std::pair<int64_t, int64_t>> orig_rest_range;   // per level [begin, end)
    std::vector<double> asm_rest_elems;             // per level: entries the level path assembles
    std::vector<double> small_rest_flops, small_rest_elems;   // per level: the level path's small fronts

    explicit DevPlan(std::shared_ptr<const Plan> p);
    void build_subtrees(const Plan &pl);
    ~DevPlan();
    DevPlan(const DevPlan &) = delete;
    DevPlan &operator=(const DevPlan &) = delete;
};

// Numeric factors of nfreq matrices with the same pattern.
template <typename T>
struct Factors {
    std::shared_ptr<DevPlan> plan;
    int nfreq = 0;
    int64_t fstride = 0;            // arena elements per frequency
    cx<T> *arena = nullptr;         // fronts, column-major, factors in place
    int *piv = nullptr;             // nfreq * n local row swaps
    int *rowperm = nullptr;         // nfreq * n: local source row of each pivot position after all swaps
    int *sing = nullptr;            // nfreq flags: zero pivot met
    int *gsrc = nullptr;            // nfreq * fv_rows * 3: gather sources after pivoting (forward sweep of M)
    cx<double> *vals = nullptr;     // nfreq * nnz FP64 values of M (CSC order) kept for refinement
    bool own_vals = false;
    int *sched = nullptr;           // task counter of the subtree launch
    // inverse squares of the single-vector path, made on its first use after a
    // factorisation (stale after any refactorisation)
    mutable cx<T> *vinv = nullptr;
    mutable bool vinv_ready = false;
    hipGraphExec_t graph = nullptr; // captured refactorisation (gfdfd_refactor_graph)
    const void *graph_src = nullptr;   // the values buffer it reads
    ~Factors();
};

// numeric factorisation of nfreq matrices whose CSC values (FP64, interleaved)
// sit at vals + f*vstride
template <typename T>
void factor_numeric(const DevPlan &dp, Factors<T> &F, const cx<double> *vals, int64_t vstride, hipStream_t st);

// Panel solve of op(M_f) X = B for one frequency; B and X are column-major
// n x nrhs interleaved complex of type TIO (may alias).  sys: 0 M, 1 M^H, 2 M^T.
template <typename T, typename TIO>
void solve_panel(const Factors<T> &F, int ifreq, int sys, int nrhs, const cx<TIO> *B, cx<TIO> *X, hipStream_t st);

// r = b - op(M) x in FP64 (CSC values vals of M); optional per-column norms ||r||_2^2 and ||b||_2^2.
void residual_fp64(const DevPlan &dp, const cx<double> *vals, int sys, int nrhs, const cx<double> *B,
                   const cx<double> *X, cx<double> *R, hipStream_t st);
void column_norms2(int n, int nrhs, const cx<double> *V, double *out, hipStream_t st);
void axpy_fp64(int64_t count, const cx<double> *D, cx<double> *X, hipStream_t st);          // X += D
void cast_c128_to_c64(int64_t count, const cx<double> *a, cx<float> *b, hipStream_t st);
void cast_c64_to_c128(int64_t count, const cx<float> *a, cx<double> *b, hipStream_t st);

// panel-solve kernel selection: widest fused right-hand-side tile (16/32/64),
// the largest front order of a multi-block level taking the fused kernels
// (negative: no limit beyond LDS), and the layout of the panel refinement
// (0: elimination order, 1: user layout with a transpose per correction)
void set_solve_tuning(int rw_cap, int fused_mmax, int refine_layout);
// factorisation kernel selection: panels needing more LDS than panel_lds_cap
// bytes (0: the hardware limit) take the global-memory panel kernel; the small
// subtrees take the persistent launch when persistent != 0, else the level path
// panel_lds_cap: panels needing more LDS take the global-memory panel kernel
// (0: the hardware limit); persistent: the small subtrees in one launch;
// deep_min_m: levels whose largest blocked front has order >= deep_min_m take
// the two-panel (depth-64) schedule, the others one panel at a time
void set_factor_tuning(int panel_lds_cap, int persistent, int deep_min_m);
// frees the calling thread's cached solve work buffers
void release_solve_workspaces();

// Solve of 1..4 right-hand sides (user layout, FP64) whose work buffers live in
// a file-scope device array of kNoAllocScratchBytes: no device allocation.
// The grid must fit: noalloc_scratch_bytes(plan, sizeof(cx<T>)) <= the array.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
size_t noalloc_scratch_bytes(const Plan &pl, size_t entry_bytes);
template <typename T>
void solve_noalloc(const Factors<T> &F, int ifreq, int sys, int nrhs, const cx<double> *B, cx<double> *X,
                   hipStream_t st);

struct RefineStats {
    int iterations = 0;
    double max_rel_residual = 0;   // max over columns of ||b - M x|| / ||b||
};

// Batched solve: launch slot z solves frequency freqs[z] on panel slot slots[z]
// of B and X (consecutive column-major n x nrhs panels).  accumulate: X += the
// solution instead of X = (refinement corrections); B and X must not alias then.
template <typename T, typename TIO>
void solve_batch(const Factors<T> &F, const std::vector<int> &freqs, const std::vector<int> &slots, int nslot_buf,
                 int sys, int nrhs, const cx<TIO> *B, cx<TIO> *X, hipStream_t st, bool accumulate = false);

// FP64 iterative refinement around the factors F (any precision):
// x_0 = solve(b); repeat r = b - M x, x += solve(r) until max rel residual <= tol.
template <typename T>
RefineStats solve_refined(const Factors<T> &F, int ifreq, int sys, int nrhs, const cx<double> *B, cx<double> *X,
                          double tol, int max_iter, hipStream_t st);
// the same for the frequencies freqs[z] on panel slots z of B and X, solved together
template <typename T>
std::vector<RefineStats> solve_refined_batch(const Factors<T> &F, const std::vector<int> &freqs, int sys, int nrhs,
                                             const cx<double> *B, cx<double> *X, double tol, int max_iter,
                                             hipStream_t st);

}  // namespace gfdfd
