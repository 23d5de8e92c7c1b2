// Internal state of an hpq context: geometry, slice-store layout, DST plans,
// multigrid hierarchy and scratch, all resident in HBM for the life of the context.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hpq.h"

namespace hpq {

#define HPQ_HIP(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct HipError : std::exception {
    std::string msg;
    explicit HipError(std::string m) : msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};
struct HpqError : std::exception {
    int code;
    std::string msg;
    HpqError(int c, std::string m) : code(c), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

// Slice geometry. Positions follow GetPosOffset (src/fields/Fields.H:72-77).
struct Geom {
    int nx = 0, ny = 0, nz = 0, ng = 2;
    int ldx = 0, ldy = 0;
    long plane = 0;
    double xlo = 0, xhi = 0, ylo = 0, yhi = 0, zlo = 0, zhi = 0;
    double dx = 0, dy = 0, dz = 0, dxi = 0, dyi = 0, dzi = 0;
    double xoff = 0, yoff = 0;
    __host__ __device__ long idx(int i, int j) const { return long(j + ng) * ldx + (i + ng); }
};

// Component indices of the slice store (src/fields/Fields.cpp:70-165).
struct Layout {
    int ncomp = 0;
    int next_jx_beam = -1, next_jy_beam = -1;
    int chi = -1, Sy = -1, Sx = -1, ExmBy = -1, EypBx = -1, Ez = -1, Bx = -1, By = -1, Bz = -1,
        Psi = -1, jx_beam = -1, jy_beam = -1, jz_beam = -1, jx = -1, jy = -1, rhomjz = -1, rho = -1;
    int prev_jx_beam = -1, prev_jy_beam = -1;
    int ions_rhomjz = -1;
    int next_jx = -1, next_jy = -1, jz = -1;
    int prev_Bx = -1, prev_By = -1, prev_jx = -1, prev_jy = -1;
    int pciter_Bx = -1, pciter_By = -1, pcprev_Bx = -1, pcprev_By = -1;
    // WhichSlice::Salame (src/fields/Fields.cpp:111-115)
    int sal_Ez_target = -1, sal_Ez_no_salame = -1, sal_Ez = -1, sal_jx = -1, sal_jy = -1, sal_jz_beam = -1,
        sal_Bx = -1, sal_By = -1, sal_Sy = -1, sal_Sx = -1, sal_Sy_back = -1, sal_Sx_back = -1;
};

// Batched 1-D DST-I (RODFT00) along contiguous rows of length n.
struct DstPlan {
    int n = 0;          // transform length
    int N = 0;          // n + 1
    int M = 0;          // FFT length 2N (power-of-two path)
    int logM = 0;
    bool fft = false;   // power-of-two FFT in LDS, else direct sum with a sine table
    double2* tw = nullptr;    // M/2 twiddles exp(-2 pi i k / M)
    double* sintab = nullptr; // 2N entries sin(pi m / N) (direct path)
};

// Dirichlet Poisson solver on one slice (FFTPoissonSolverDirichlet*).
struct Poisson {
    int nx = 0, ny = 0;
    DstPlan px, py;
    double* eigT = nullptr;   // [i][j] eigenvalue matrix with the RODFT00 normalisation
    double inv_norm = 0;      // 1 / the normalisation factor in eigT
    double* chi_part = nullptr;       // HPQ_MG_GUESS_DST: sums of chi over each row
    double* bufA = nullptr;   // 3 * nx * ny
    double* bufB = nullptr;
    // sine-matrix GEMM path (HPQ_POISSON_BF16X3): sizes padded to 64, S split into BF16
    // hi/lo planes [j][k], FP32 intermediates 3 * npx * npy
    int npx = 0, npy = 0;
    __bf16 *sx_hi = nullptr, *sx_lo = nullptr, *sy_hi = nullptr, *sy_lo = nullptr;
    float *gA = nullptr, *gB = nullptr;
    float* gS = nullptr;      // staged right-hand sides [f][npy][npx], zero padding
};

// One multigrid level; arrays cover the level box plus one ghost ring.
struct MGLevel {
    bool cc = true;
    int ilo = 0, ihi = 0, jlo = 0, jhi = 0;
    int vilo = 0, vihi = 0, vjlo = 0, vjhi = 0;
    int ldx = 0, ldy = 0;
    long cs = 0;   // component stride
    double dx = 0, dy = 0;
    double *acf = nullptr, *res = nullptr, *cor = nullptr, *rescor = nullptr;
};

// Convergence state of one solve, kept on the device so V-cycles can be enqueued without
// a host round trip per cycle; max norms are accumulated as the bit patterns of
// non-negative doubles (their integer order is the numeric order). Two of them alternate
// between solves: the kernel that ends a solve clears the other one for the next solve.
struct MGCtl {
    int done = 0;
    int iters = 0;
    int status = 0;
    int pad = 0;
    double max_norm = 0;
    double res_target = 0;
};

// What the device publishes to the host after every V-cycle, in one 64-bit word of pinned
// coherent host memory: seq (bits 0-31), iters (32-55), status (56-62), done (63).
struct MGMirror {
    unsigned long long word;
};

struct MultiGrid {
    bool cc = true;
    int max_level = 0;
    int coarse_begin = 0;   // first level handled by the single-workgroup coarse kernel
    bool coarse_in_lds = false;
    size_t coarse_lds_bytes = 0;
    int last_iters = 1;     // V-cycles of the previous solve: first batch size of the next
    bool last_zero = false; // the previous solve converged without a V-cycle
    double dx = 0, dy = 0;
    std::vector<MGLevel> lev;
    alignas(16) unsigned char coarse_args[2048];   // kernel argument block of the coarse kernel
    double* coarse_work = nullptr;   // coarse-level arrays when they do not fit in LDS
    double* msol = nullptr;          // level-0 initial guess / odd iterates (copy path, 2 comps)
    double* mrhs = nullptr;          // level-0 right-hand side (copy path, 2 comps)
    MGCtl* d_ctl = nullptr;          // two control blocks, used alternately
    double* d_norms = nullptr;       // max |residual|, max |rhs| of every level-0 workgroup
    int ctl_pick = 0;
    bool ctl_dirty = false;          // a solve ended early: clear both before the next one
    MGMirror* h_mirror = nullptr;    // pinned coherent host memory
    MGMirror* d_mirror = nullptr;    // its device address
    int seq = 0;
    // HPQ_MG_GUESS_LINEAR: Bx, By of the slice before the last one solved (2 planes)
    double* hist = nullptr;
    int hist_slice = -1;             // the slice whose Bx, By hist holds; -1: none
};

struct Capture {
    int comp;
    double* dst;
};

// Tile sort of the plasma for deposition (bins.hip): scratch of the counting sort and the
// permutation the deposition kernels walk.
// Deposition tiles are 32 x 8 cells (256 cells, a 2 x 2-ppc chunk of 1024 particles): in
// index order a tile's particles fill a wavefront's lanes one row of 32 cells at a time, so the
// 32 lanes of a half-wavefront touch 32 consecutive LDS doubles.
constexpr int DEPOSIT_TILE_X = 32, DEPOSIT_TILE_Y = 8;
struct Bins {
    int32_t* keys = nullptr;
    long keys_cap = 0;
    int32_t* counts = nullptr;   // (bin, chunk) counts, then their exclusive scan
    long counts_cap = 0;
    int32_t* tmp = nullptr;      // block sums of the scan
    long tmp_cap = 0;
    int32_t* perm = nullptr;     // deposition order
    long perm_cap = 0;
    long perm_n = -1;
    bool stale = true;           // re-sort before the next deposition
    int calls = 0;               // plasma depositions since the last re-sort
    double* reorder = nullptr;   // staging of the plasma arrays while they are reordered
    long reorder_cap = 0;
    bool reordered = false;      // the plasma was reordered since the step began
};

// Per-start-cell particle lists of the ordered deposition (gather.hip): list offsets over a
// grid of base cells, the particles of every list by increasing index, and the per-particle
// values the contributions need. valid/base/x/n name the plasma state the lists describe, so
// the explicit deposition can reuse the lists of the current deposition on the same state.
struct CellLists {
    int bxlo = 0, bylo = 0, wx = 0, wy = 0;
    int T = 32, ntx = 1, nty = 1;
    int32_t *off = nullptr, *tmp = nullptr, *key = nullptr, *slot = nullptr, *unsorted = nullptr, *perm = nullptr;
    long off_cap = 0, tmp_cap = 0, key_cap = 0, slot_cap = 0, unsorted_cap = 0, perm_cap = 0;
    double* rec = nullptr;
    long rec_cap = 0;
    int32_t* ovf = nullptr;      // overflow lists of the staged gathers
    long ovf_cap = 0;
    bool explicit_ovf_clear = false;   // the explicit deposition's overflow count is zero
    // the beam's per-cell contribution sequences: segment offsets per plane cell, their scan
    // scratch, the contributions (3 arrays), the cells summed by a wavefront each
    int32_t *seg = nullptr, *segtmp = nullptr, *longc = nullptr;
    long seg_cap = 0, segtmp_cap = 0, longc_cap = 0;
    double* contrib = nullptr;
    long contrib_cap = 0;
    bool valid = false;
    int base = -1;
    const double* x = nullptr;
    long n = 0;
};

// Scratch of the plasma injection: per-cell particle counts and their exclusive scan.
struct InitScratch {
    int32_t* off = nullptr;
    int32_t* tmp = nullptr;
    long cap = 0;
};

// Plasma SoA copy (11 FP64 arrays in PlasmaIdx order + valid) for checkpoints.
constexpr int PLASMA_REALS = 11;
struct PlasmaBuf {
    double* r[PLASMA_REALS] = {};
    int32_t* valid = nullptr;
    long n = 0;
};

// State of the adjoint slice loop: checkpoints of the forward pass (slice store and plasma
// every k slices), the per-slice saves of one recomputed segment, the adjoint planes and
// two sets of plasma adjoints (the state entering and leaving a slice).
struct AdjointState {
    int k = 0, nck = 0;
    long np = 0;
    std::vector<double*> ck_store, sl_store;
    std::vector<PlasmaBuf> ck_plasma, sl_plasma;
    double* store_adj = nullptr;
    double* padj[2][8] = {};
    double* mu = nullptr;        // Bx/By adjoint solution of the previous slice (warm start)
    double* push_g = nullptr;    // adjoints of the six gathered push fields per particle
    long push_g_cap = 0;
    bool ready = false;
};

// Slipped beam particles between consecutive slice pushes: two index lists and their counts.
struct BeamSlip {
    int32_t* idx[2] = {nullptr, nullptr};
    int32_t* cnt = nullptr;      // device, two counters
    long cap = 0;
    int cur = 0;                 // list holding the particles slipped into the next slice pushed
};

// Event pairs around the launches of one kernel family (HPQ_TRACE_*).
struct Trace {
    bool on = false;
    std::vector<hipEvent_t> start, stop;
    size_t used = 0;
};

}  // namespace hpq

struct hpq_ctx {
    hpq_params p{};
    hpq::Geom g;
    hpq::Layout L;
    hipStream_t stream = nullptr;
    int device = -1;                 // HIP device current when the context was created
    int inj_ppc = 1;                 // particles per cell of the last hpq_init_plasma lattice
    hpq::Poisson ps;
    hpq::MultiGrid mg;
    hpq::Bins bins;
    hpq::CellLists cl_plasma;        // ordered deposition of the plasma (main stream)
    hpq::CellLists cl_beam;          // ordered deposition of the beam (either stream)
    hpq::InitScratch init;
    double* relb_rows = nullptr;     // row sums of ComputeRelBFieldError (2 * ny)
    long relb_cap = 0;
    hpq::AdjointState adj;
    hpq::BeamSlip slip;
    int* d_count = nullptr;          // QSA counter
    int* h_count = nullptr;          // pinned
    double* d_red = nullptr;         // reduction scratch (8 doubles)
    double* h_red = nullptr;         // pinned
    double* checksums = nullptr;     // caller-owned device accumulator (may be null)
    double* diag_partial = nullptr;  // per-workgroup diagnostic sums not yet in checksums
    int* diag_comp_of = nullptr;
    bool diag_pending = false;
    hpq_insitu insitu{};             // caller-owned in-situ diagnostic arrays (all NULL: off)
    hpq_salame salame{};             // SALAME settings (mask NULL: off)
    int sal_last_slice = -1;         // Hipace::m_salame_last_slice
    bool sal_overloaded = false;     // Hipace::m_salame_overloaded
    double sal_zeta_initial = 0.0;   // Hipace::m_salame_zeta_initial
    double* sal_dev = nullptr;       // device: 4 sums of SalameGetW, particle count
    double* sal_host = nullptr;      // pinned copy
    // second stream of the chained slice loop: beam deposition and the end-of-slice pass run
    // beside the plasma deposition / Poisson solves and the plasma push
    hipStream_t side = nullptr;
    hipEvent_t ev_start = nullptr, ev_beam = nullptr, ev_fields = nullptr, ev_end = nullptr;
    double* insitu_partial = nullptr;  // per-workgroup in-situ sums of one slice
    long insitu_cap = 0;               // workgroups insitu_partial holds
    std::vector<hpq::Capture> captures;
    double stats[4] = {0, 0, 0, 0};
    hpq::Trace trace[HPQ_TRACE_N];
    std::string err;
};

namespace hpq {

// kernels and helpers implemented in the .hip translation units
void layout_init(hpq_ctx* c);
void poisson_init(hpq_ctx* c);
void poisson_free(hpq_ctx* c);
void poisson_solve_staged(hpq_ctx* c, int nfields, double* const* lhs_planes);
void dst_mfma_init(hpq_ctx* c);
void dst_mfma_free(hpq_ctx* c);
void poisson_solve_mfma(hpq_ctx* c, int nfields, double* const* lhs_planes);
void stage_rhs(hpq_ctx* c, const double* src);
void build_rhs_psi_ez_bz(hpq_ctx* c, const double* store);
void build_rhs_bxby(hpq_ctx* c, const double* store);
void exmby_eypbx(hpq_ctx* c, double* store);
void mg_init(hpq_ctx* c);
void mg_free(hpq_ctx* c);
// The multigrid's initial guess for the Bx, By of slice islice (two planes from bxby, one
// plane apart) under HPQ_MG_GUESS_LINEAR: 2 B(islice + 1) - B(islice + 2) when the history
// holds slice islice + 2 and bxby slice islice + 1; otherwise bxby is kept and recorded.
void mg_guess_linear(hpq_ctx* c, double* bxby, int islice);
// HPQ_MG_GUESS_DST: iters corrections (Bx, By) += (Lap - mean chi)^-1 [(Sy, Sx) - (Lap - chi) (Bx, By)]
// by the FFT DST solve, starting from the store's Bx, By (the previous slice's solution)
void bxby_guess_dst(hpq_ctx* c, double* store, int iters);
int mg_solve1(hpq_ctx* c, double* sol, const double* rhs, const double* acf, long comp_stride, double tol_rel,
              double tol_abs, int maxiter, int* iters);
void deposit_current(hpq_ctx* c, const hpq_plasma* p, double* store, const int32_t* comps, bool ions);
void explicit_deposition(hpq_ctx* c, const hpq_plasma* p, double* store);
void deposit_beam(hpq_ctx* c, const hpq_beam* b, int slice, double* store, const int32_t* comps);
void advance_plasma(hpq_ctx* c, const hpq_plasma* p, const double* store, bool temp_slice);
void plane_fill(hpq_ctx* c, double* store, const int* comps, int ncomps, double v);
void plane_copy(hpq_ctx* c, double* store, const int* dst, const int* src, int ncomps);
void plane_add(hpq_ctx* c, double* store, const int* dst, const int* src, int ncomps);
void plane_lincomb(hpq_ctx* c, double* store, int dst, double a, int sa, double b, int sb);
void sxsy_from_beam(hpq_ctx* c, double* store);
void captures(hpq_ctx* c, const double* store, int islice);
void slice_end(hpq_ctx* c, double* store, int islice, const int* shift_dst, const int* shift_src, int nshift,
               const int* zero, int nzero);
void diag_finalize(hpq_ctx* c);
constexpr int INSITU_STRIDE = 24;  // doubles per workgroup in insitu_partial
void insitu_reserve(hpq_ctx* c);
double* insitu_scratch(hpq_ctx* c, long nblk);
void insitu_finalize(hpq_ctx* c, int nblk, int nreal, bool count, double scale, uint32_t norm_mask, double* rdata,
                     double* sum, int32_t* idata, int32_t* isum, int islice);
void insitu_plasma(hpq_ctx* c, const hpq_plasma* p, int islice);
void insitu_beam(hpq_ctx* c, const hpq_beam* b, int islice);
void insitu_free(hpq_ctx* c);
void deposit_beam_select(hpq_ctx* c, const hpq_beam* b, int slice, double* store, const int32_t* comps,
                         const int32_t* select, int select_value);
void salame_alloc(hpq_ctx* c);
void salame_free(hpq_ctx* c);
long salame_count(hpq_ctx* c, const hpq_beam* b, int slice);
void salame_sxsy(hpq_ctx* c, double* store);
void salame_rhs_ez(hpq_ctx* c, const double* store);
void salame_only_advance(hpq_ctx* c, const hpq_plasma* p, double* store);
void salame_jxjy_from_b(hpq_ctx* c, double* store);
void salame_sums(hpq_ctx* c, const double* store, double out[4]);
void salame_multiply(hpq_ctx* c, const hpq_beam* b, int slice, double W);
double rel_b_error(hpq_ctx* c, const double* store, int bx, int by, int bxi, int byi);
void init_plasma(hpq_ctx* c, const hpq_plasma_profile* prof, const hpq_plasma* p);
void init_scratch_free(hpq_ctx* c);
int64_t plasma_count(hpq_ctx* c, const hpq_plasma_profile* prof);
void bin_tiles(hpq_ctx* c, const hpq_plasma* p, int tile, int32_t* perm, int32_t* offsets);
void bin_tiles_rect(hpq_ctx* c, const hpq_plasma* p, int tile, int tile_y, int split, int32_t* perm,
                    int32_t* offsets);
// plasmas.reorder_period: the plasma arrays reordered in place by deposition tile
void reorder_plasma(hpq_ctx* c, const hpq_plasma* p);
// the depositions sum in the reference CPU build's order (FP64 path, ordered deposition)
inline bool ordered_deposition(const hpq_ctx* c) {
    return c->p.deposition == HPQ_DEPOSIT_ORDERED && c->p.precision == HPQ_PREC_FP64;
}
// the atomic depositions walk the plasma in index order (reordered in place every
// reorder_period slices) instead of through the tile permutation
inline bool plasma_reordered(const hpq_ctx* c) { return c->p.reorder_period > 0 && !ordered_deposition(c); }
const int32_t* deposit_order(hpq_ctx* c, const hpq_plasma* p);
void bins_free(hpq_ctx* c);
void sort_keys(hpq_ctx* c, long n, int nbins, int32_t* perm, int32_t* offsets);
// tile binning with its work arrays in a file-scope device array (no allocation)
void bins_noalloc_setup(int device);
void bin_tiles_noalloc(hpq_ctx* c, const hpq_plasma* p, int tile, int32_t* perm, int32_t* offsets);
// exclusive scan of a[0..n) in place; tmp: scan_tmp_size(n) entries
void scan_exclusive(hipStream_t s, int32_t* a, long n, int32_t* tmp);
long scan_tmp_size(long n);
void deposit_current_ordered(hpq_ctx* c, const hpq_plasma* p, double* store, const int32_t* comps, bool ions);
void explicit_deposition_ordered(hpq_ctx* c, const hpq_plasma* p, double* store);
void deposit_beam_ordered(hpq_ctx* c, const hpq_beam* b, int slice, double* store, const int32_t* comps,
                          const int32_t* select, int select_value);
void cell_lists_free(CellLists& L);
void sort_beam_slices(hpq_ctx* c, hpq_beam_state* b, int64_t* off_out);
void deposit_vjp(hpq_ctx* c, const hpq_plasma* p, const double* adj, const int32_t* comps, hpq_plasma_adj* a);
void explicit_deposition_vjp(hpq_ctx* c, const hpq_plasma* p, const double* store, double* adj, hpq_plasma_adj* a);
void advance_plasma_vjp(hpq_ctx* c, const hpq_plasma* p, const double* store, const hpq_plasma_adj* out,
                        hpq_plasma_adj* in, double* adj);
void adj_poisson(hpq_ctx* c, double* adj);
void adj_chi(hpq_ctx* c, const double* store, double* adj);
void adj_seed(hpq_ctx* c, const double* seed, double* adj);
void adj_beam_weights(hpq_ctx* c, const hpq_beam* b, int slice, double* adj, const double* jzb_seed,
                      double* w_adj);
void adjoint_free(hpq_ctx* c);
void beam_begin_step(hpq_ctx* c, hpq_beam_state* b);
void advance_beam(hpq_ctx* c, hpq_beam_state* b, int slice, const double* store, double dt_step, int n_subcycles,
                  bool do_z_push);
void beam_slip_free(hpq_ctx* c);

inline dim3 grid1d(long n, int block) { return dim3(unsigned((n + block - 1) / block)); }

// The 8 XCDs of an MI355X receive the workgroups of a 1-D grid round robin (workgroup b on
// XCD b % 8), and each XCD has its own L2. Maps the physical workgroup index onto a logical
// one such that every XCD works on one contiguous range of logical workgroups (XCD x < r
// takes q + 1 of them, the others q, with n = 8 q + r), so neighbouring particles and the
// field rows they read stay in one L2. A bijection of [0, n).
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block(int b, int n) {
    const int q = n / NUM_XCD, r = n % NUM_XCD;
    const int x = b % NUM_XCD, i = b / NUM_XCD;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// calls f(std::integral_constant<int, O>) for the run-time shape order O
template <class F>
void dispatch_order(int order, F&& f) {
    switch (order) {
        case 0: f(std::integral_constant<int, 0>{}); break;
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        default: throw HpqError(HPQ_E_UNSUPPORTED, "depos_order_xy must be 0..3");
    }
}

// Records HIP events around the launches in its scope when tracing of `kind` is on.
class TraceScope {
public:
    TraceScope(hpq_ctx* c, int kind);
    ~TraceScope();
    TraceScope(const TraceScope&) = delete;
    TraceScope& operator=(const TraceScope&) = delete;

private:
    hpq_ctx* c_;
    int kind_;
    size_t slot_;
};

}  // namespace hpq
