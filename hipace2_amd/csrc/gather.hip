// Deposition in the reference CPU build's summation order, on gfx950, without atomics.
//
// The CPU build deposits with SharedMemoryDeposition's tiled loop
// (src/particles/deposition/DepositionUtil.H:204-253; tiling on, 32-cell tiles under OpenMP,
// src/Hipace.H:235-240): particles are binned by the tile of their lowest deposited cell
// (start cell), the tiles are visited in four colours (even/odd tile index in x, then in y),
// and inside a tile the particles go in bin order, which is index order (amrex::DenseBins is a
// stable counting sort). Tiles of one colour never share a cell, so every cell of the slice
// receives its contributions in one fixed sequence: the up to four tiles that reach it in
// colour order, and inside each tile the particles by index.
//
// Here every cell sums its own contributions in that sequence (one thread per cell, the
// running sum starts from the plane's value and each contribution is added in turn), so the
// planes come out bit for bit as the CPU build's loop leaves them, and no two threads write
// the same cell. The particles that reach a cell are those whose start cell lies in the
// S x S window below it (S = stencil width); they are found through per-start-cell lists:
//   keys    the start cell of every particle (QSA violators are removed here, as the deposit
//           functor would remove them) and its slot in that cell's list (one atomic per
//           distinct cell of a wavefront, slots in lane order);
//   count   particles per start cell, exclusive scan -> list offsets;
//   scatter particles into their lists (atomic slots, any order);
//   order   each list sorted by particle index (rank by counting inside the list); at its
//           list position the particle's record is written: the per-particle values every
//           contribution needs (q w, u/psi, gamma/psi, 1/psi, the cell-scaled position), so
//           the walk along a list reads consecutive memory;
//   gather  per cell, for each colour in turn, the lists of the window cells in that colour's
//           tile merged by particle index; each particle's contribution is the reference
//           expression evaluated for this cell's stencil offset. Workgroups are mapped to the
//           XCDs in contiguous bands of cell blocks, so the lists and records neighbouring
//           blocks share stay in one L2.
// A start-cell list is keyed by the floor of the cell-scaled position the shape functions
// take their cell from (floor(xmid) or floor(xmid + 1/2)); a stencil's start cell is that
// floor plus a constant, so the explicit Sx/Sy deposition of the default derivative type
// reuses the lists the current deposition built on the same particles.
#include <climits>
#include <cmath>

#include "internal.hpp"
#include "shape.hpp"

namespace hpq {

namespace {


// start cell = floor(xmid + 0.5 * BASE) + DELTA, for the order-O shape factor and for the
// type-T derivative shape factor (ShapeFactors.H: compute_single_shape_factor,
// single_derivative_shape_factor)
template <int O> constexpr int shape_base() { return (O == 0 || O == 2) ? 1 : 0; }
template <int O> constexpr int shape_delta() { return O >= 2 ? -1 : 0; }
template <int T, int O> constexpr int dshape_base() {
    return T == 0 ? shape_base<O>() : (T == 1 ? ((O == 1 || O == 3) ? 1 : 0) : ((O == 0 || O == 2) ? 1 : 0));
}
template <int T, int O> constexpr int dshape_delta() {
    return T == 0 ? shape_delta<O>() : (T == 1 ? (O == 0 ? 0 : (O == 3 ? -2 : -1)) : (O >= 2 ? -2 : -1));
}

// lists of one particle set keyed by base cell: base cells [blo, blo + w) in each direction;
// tiles of the deposition order: T cells (hipace.tile_size), ntx x nty over the guarded box
struct CLView {
    int bxlo, bylo, wx, wy;
    int T, ntx, nty;
    const int32_t* __restrict__ off;    // wx * wy + 1 list offsets
    const int32_t* __restrict__ perm;   // particle indices, list by list, increasing in a list
};

// the tile of start cell s along one direction (C division truncates toward zero, as the
// reference's (cell - lo) / tile does; clamped to the tile grid)
__device__ __forceinline__ int dep_tile(int s, int ng, int T, int nt) { return min(max((s + ng) / T, 0), nt - 1); }

// A list source for the walk over a cell's window: list (kx, ky) of the window covers entries
// [bound(kx, ky), bound(kx + 1, ky)) of the source (a row of the window is contiguous);
// idx(h) is the particle index of entry h, pos(h, ky) the position of its record.
struct GlobalLists {
    CLView cl;
    int row0, col0;   // list-grid row and column of window list (0, 0)
    __device__ __forceinline__ int bound(int kx, int ky) const { return cl.off[(row0 + ky) * cl.wx + col0 + kx]; }
    __device__ __forceinline__ int idx(int h) const { return cl.perm[h]; }
    __device__ __forceinline__ int pos(int h, int) const { return h; }
};

// Calls f(pos, ix, iy) for every particle of the lists whose start cell lies in the S x S
// window below cell (i, j), in the CPU build's order: colour by colour, and inside the window
// part of one colour by increasing particle index (a merge of the lists). pos is the
// particle's record position, (ix, iy) its stencil offset that lands on (i, j).
template <int S, class Src, class F>
__device__ __forceinline__ void walk_ordered(const Geom& g, const CLView& cl, int i, int j, const Src& src, F&& f) {
    const int x0 = i - S + 1, y0 = j - S + 1;
    int h[S * S], e[S * S];
#pragma unroll
    for (int ky = 0; ky < S; ++ky) {
        int o[S + 1];
#pragma unroll
        for (int kx = 0; kx <= S; ++kx) o[kx] = src.bound(kx, ky);
#pragma unroll
        for (int kx = 0; kx < S; ++kx) {
            h[ky * S + kx] = o[kx];
            e[ky * S + kx] = o[kx + 1];
        }
    }
    // window split at tile boundaries: columns [x0, xs) in tile txa, [xs, i] in tile txb
    // (S <= T: a window spans at most two tiles in each direction)
    const int txa = dep_tile(x0, g.ng, cl.T, cl.ntx), txb = dep_tile(i, g.ng, cl.T, cl.ntx);
    const int tya = dep_tile(y0, g.ng, cl.T, cl.nty), tyb = dep_tile(j, g.ng, cl.T, cl.nty);
    const int xs = txa == txb ? i + 1 : txb * cl.T - g.ng;
    const int ys = tya == tyb ? j + 1 : tyb * cl.T - g.ng;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        // the x part and the y part of this colour (empty when no window tile has it)
        int ax, bx, ay, by;
        if ((txa & 1) == (c >> 1)) { ax = x0; bx = xs - 1; }
        else if ((txb & 1) == (c >> 1)) { ax = xs; bx = i; }
        else { ax = 1; bx = 0; }
        if ((tya & 1) == (c & 1)) { ay = y0; by = ys - 1; }
        else if ((tyb & 1) == (c & 1)) { ay = ys; by = j; }
        else { ay = 1; by = 0; }
        if (ax > bx || ay > by) continue;
        int v[S * S];
#pragma unroll
        for (int k = 0; k < S * S; ++k) {
            const int sx = x0 + k % S, sy = y0 + k / S;
            const bool act = sx >= ax && sx <= bx && sy >= ay && sy <= by && h[k] < e[k];
            v[k] = act ? src.idx(h[k]) : INT_MAX;
        }
        while (true) {
            int m1 = INT_MAX, m2 = INT_MAX, km = -1;
#pragma unroll
            for (int k = 0; k < S * S; ++k) {
                if (v[k] < m1) {
                    m2 = m1;
                    m1 = v[k];
                    km = k;
                } else if (v[k] < m2) {
                    m2 = v[k];
                }
            }
            if (km < 0) break;
            int hc = 0, ec = 0;
#pragma unroll
            for (int k = 0; k < S * S; ++k)
                if (k == km) { hc = h[k]; ec = e[k]; }
            const int ix = S - 1 - km % S, iy = S - 1 - km / S;
            const int kyl = km / S;
            // the list holding the smallest index goes on while it stays below every other head
            int p = m1;
            while (true) {
                f(src.pos(hc, kyl), ix, iy);
                ++hc;
                p = hc < ec ? src.idx(hc) : INT_MAX;
                if (p >= m2) break;
            }
#pragma unroll
            for (int k = 0; k < S * S; ++k)
                if (k == km) { h[k] = hc; v[k] = p; }
        }
    }
}

// ---- block staging -------------------------------------------------------------------------
// A gather workgroup owns GS x GS cells. Their windows together cover (GS + S - 1)^2 lists,
// GS + S - 1 rows of consecutive lists, each row one contiguous range of the list entries:
// the workgroup copies the list bounds and the particle indices of those rows into LDS, so
// every cell merges its window from LDS. A block whose rows hold more entries than the
// staging area (densities far above the lattice's, the core of a beam) is listed instead,
// and a second launch of the same kernel merges the listed blocks from global memory.
constexpr int GS = 16;             // cells per side of a gather block
constexpr int GS_MAXW = GS + 5;    // widest window side (stencils of at most 6 cells)
constexpr int GS_CAP = 6144;       // staged list entries of the plasma gathers (24 KiB)
constexpr int GS_CAP_BEAM = 24576; // of the beam gather (96 KiB): a beam's core is dense
template <int CAP>
struct StagedLists {
    int off[GS_MAXW][GS_MAXW + 1];   // window list bounds, in staged entries
    int base[GS_MAXW];               // record position = staged entry + base[row]
    int total;                       // staged entries (more than CAP: not staged)
    int idx[CAP];                    // particle index of every staged entry
};
template <int CAP>
struct StagedView {
    const StagedLists<CAP>& s;
    int li, lj;   // the cell's column and row in the block
    __device__ __forceinline__ int bound(int kx, int ky) const { return s.off[lj + ky][li + kx]; }
    __device__ __forceinline__ int idx(int h) const { return s.idx[h]; }
    __device__ __forceinline__ int pos(int h, int ky) const { return h + s.base[lj + ky]; }
};

__device__ __forceinline__ int gather_nbx(const Geom& g) { return (g.ldx + GS - 1) / GS; }
__device__ __forceinline__ int gather_nby(const Geom& g) { return (g.ldy + GS - 1) / GS; }

// stages the rows of lists below block (i0, j0) for stencils of S cells starting DELTA from
// the base cell; every thread of the workgroup calls it
template <int S, int DELTA, int CAP>
__device__ __forceinline__ void stage_lists(const CLView& cl, int i0, int j0, StagedLists<CAP>& sl) {
    constexpr int NW = GS + S - 1;
    static_assert(NW <= GS_MAXW, "window wider than the staging area");
    const int col0 = i0 - S + 1 - DELTA - cl.bxlo;
    const int row0 = j0 - S + 1 - DELTA - cl.bylo;
    // list bounds; lists outside the list grid are empty
    for (int e = threadIdx.x; e < NW * (NW + 1); e += blockDim.x) {
        const int r = e / (NW + 1), c = e - r * (NW + 1);
        const int row = row0 + r;
        sl.off[r][c] = (row >= 0 && row < cl.wy) ? cl.off[row * cl.wx + min(max(col0 + c, 0), cl.wx)] : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int r = 0; r < NW; ++r) {
            const int a = sl.off[r][0];
            sl.base[r] = a - run;
            run += sl.off[r][NW] - a;
        }
        sl.total = run;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NW * (NW + 1); e += blockDim.x) {
        const int r = e / (NW + 1), c = e - r * (NW + 1);
        sl.off[r][c] -= sl.base[r];
    }
    __syncthreads();
    const int total = sl.total;
    if (total <= CAP) {
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            // the row holding staged entry t: the last row starting at or before t
            int lo = 0, hi = NW - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sl.off[mid][0] <= t) lo = mid;
                else hi = mid - 1;
            }
            sl.idx[t] = cl.perm[t + sl.base[lo]];
        }
    }
    __syncthreads();
}

// Records are loaded one element ahead: the walk hands an element to `emit`, which issues
// the loads of its record and then adds the previous element, whose loads were issued one
// step earlier, so a record's memory latency overlaps the merge of the next element.
template <class Rec, class Load, class Add>
struct Pipelined {
    Load load;
    Add add;
    Rec r{};
    int ix = 0, iy = 0;
    bool pend = false;
    __device__ __forceinline__ void operator()(int pos, int nix, int niy) {
        const Rec nr = load(pos, nix, niy);
        if (pend) add(r, ix, iy);
        r = nr;
        ix = nix;
        iy = niy;
        pend = true;
    }
    __device__ __forceinline__ void finish() {
        if (pend) add(r, ix, iy);
        pend = false;
    }
};
template <class Rec, class Load, class Add>
__device__ __forceinline__ Pipelined<Rec, Load, Add> pipelined(Load load, Add add) {
    return Pipelined<Rec, Load, Add>{load, add};
}

// The colours (0..3, the CPU build's tile loop order) of the lists of the window of cell
// (i, j): a window of S <= T cells spans at most two tiles per direction, split at column xs
// and row ys; bit k of mask[c] is set when list k = ky * S + kx has colour c (S <= 6).
template <int S>
struct WindowColours {
    unsigned long long mask[4];   // S * S <= 36 lists
    __device__ __forceinline__ WindowColours(const Geom& g, const CLView& cl, int i, int j) {
        const int x0 = i - S + 1, y0 = j - S + 1;
        const int txa = dep_tile(x0, g.ng, cl.T, cl.ntx), txb = dep_tile(i, g.ng, cl.T, cl.ntx);
        const int tya = dep_tile(y0, g.ng, cl.T, cl.nty), tyb = dep_tile(j, g.ng, cl.T, cl.nty);
        const int xs = txa == txb ? i + 1 : txb * cl.T - g.ng;
        const int ys = tya == tyb ? j + 1 : tyb * cl.T - g.ng;
#pragma unroll
        for (int c = 0; c < 4; ++c) mask[c] = 0ull;
#pragma unroll
        for (int k = 0; k < S * S; ++k) {
            const int kx = k % S, ky = k / S;
            const int tx = x0 + kx < xs ? txa : txb, ty = y0 + ky < ys ? tya : tyb;
            const int c = ((tx & 1) << 1) | (ty & 1);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (c == q) mask[q] |= 1ull << k;
        }
    }
};

// Whether the merged order of the window is the raster order of its lists: for every colour,
// the non-empty lists of that colour, taken row by row, hold increasing particle indices
// (each list's first index above the previous list's last). The usual case: a lattice
// plasma's indices follow its cells, and particles move less than a cell per slice.
template <int S, class View>
__device__ __forceinline__ bool raster_ordered(const WindowColours<S>& wc, const View& src) {
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int last = -1;
#pragma unroll
        for (int k = 0; k < S * S; ++k) {
            const int kx = k % S, ky = k / S;
            const int a = src.bound(kx, ky), b = src.bound(kx + 1, ky);
            if (((wc.mask[c] >> k) & 1ull) && a < b) {
                ok = ok && src.idx(a) > last;
                last = src.idx(b - 1);
            }
        }
    }
    return ok;
}

// Two passes over the blocks of GS x GS cells (row by row over the guarded plane, each XCD
// taking one contiguous band of them). FAST: the workgroup's own block, its rows staged in
// LDS; when every cell's merged order is the raster order of its lists, the cells walk them
// in that order; otherwise (or when the rows do not fit in LDS) the block is appended to the
// list ovf (ovf[0] blocks in ovf[1..]) and left to the second pass. MERGE: the blockIdx.x-th
// listed block, each cell merging its lists (from LDS when they fit, from global memory when
// they do not). The fast pass keeps few registers; the merge's state stays in the second.
template <bool FAST, int S, int DELTA, int CAP>
__device__ __forceinline__ bool gather_setup(const Geom& g, const CLView& cl, StagedLists<CAP>& sl,
                                             int32_t* __restrict__ ovf, int& b, int& i0, int& j0) {
    const int nbx = gather_nbx(g);
    if constexpr (FAST) {
        b = xcd_block(int(blockIdx.x), nbx * gather_nby(g));
    } else {
        if (int(blockIdx.x) >= ovf[0]) return false;
        b = ovf[1 + blockIdx.x];
    }
    i0 = (b % nbx) * GS - g.ng;
    j0 = (b / nbx) * GS - g.ng;
    stage_lists<S, DELTA>(cl, i0, j0, sl);
    if (FAST && sl.total > CAP) {
        if (threadIdx.x == 0) ovf[1 + atomicAdd(&ovf[0], 1)] = b;
        return false;
    }
    return true;
}

// fast pass: whether every cell of the block can walk its lists in raster order; a block that
// cannot is listed for the merge pass (every thread of the workgroup calls this)
template <int S, int CAP>
__device__ __forceinline__ bool block_raster_ordered(const Geom& g, const CLView& cl, const StagedLists<CAP>& sl,
                                                     bool in, int i, int j, int li, int lj, int b,
                                                     int32_t* __restrict__ ovf) {
    const bool bad = in && !raster_ordered<S>(WindowColours<S>(g, cl, i, j), StagedView<CAP>{sl, li, lj});
    if (__syncthreads_or(bad)) {
        if (threadIdx.x == 0) ovf[1 + atomicAdd(&ovf[0], 1)] = b;
        return false;
    }
    return true;
}

// The walk of cell (i, j) = block cell (li, lj) in the CPU build's order, add(load(pos), ix, iy)
// for every contribution. FAST: the lists in raster order (stencil offsets known at compile
// time, records read ahead by unrolling). MERGE: the merge of the lists, from LDS when the
// block's rows were staged, from global memory otherwise.
template <int S, int DELTA, bool FAST, int CAP, class Load, class Add>
__device__ __forceinline__ void walk_cell(const Geom& g, const CLView& cl, const StagedLists<CAP>& sl, int i, int j,
                                          int li, int lj, Load&& load, Add&& add) {
    using Rec = decltype(load(0, 0, 0));
    const StagedView<CAP> src{sl, li, lj};
    if constexpr (FAST) {
        const WindowColours<S> wc(g, cl, i, j);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (int k = 0; k < S * S; ++k) {
                const int kx = k % S, ky = k / S;
                if (!((wc.mask[c] >> k) & 1ull)) continue;
                const int a = src.bound(kx, ky), e = src.bound(kx + 1, ky);
#pragma unroll 2
                for (int h = a; h < e; ++h) add(load(src.pos(h, ky), S - 1 - kx, S - 1 - ky), S - 1 - kx, S - 1 - ky);
            }
        }
    } else {
        auto f = pipelined<Rec>(load, add);
        if (sl.total <= CAP) {
            walk_ordered<S>(g, cl, i, j, src, f);
        } else {
            const GlobalLists gsrc{cl, j - S + 1 - DELTA - cl.bylo, i - S + 1 - DELTA - cl.bxlo};
            walk_ordered<S>(g, cl, i, j, gsrc, f);
        }
        f.finish();
    }
}

// ---- list construction --------------------------------------------------------------------

// The deposit functor's per-particle values of the current and explicit deposition
// (PlasmaDepositCurrent.cpp:182-243, ExplicitDeposition.cpp:174-205; the two evaluate them
// with the same operations): [0] q w invvol, [1] vx/c = ux/psi, [2] vy/c, [3] gamma/psi,
// [4] 1/psi, [5] xmid, [6] ymid.
constexpr int PLASMA_NREC = 7;
struct PlasmaVals {
    double q_invvol, vx_c, vy_c, gamma_psi, psi_inv, xmid, ymid;
};
__device__ __forceinline__ PlasmaVals plasma_vals(const Geom& g, const hpq_plasma& p, long ip, double charge_invvol) {
    const double clightinv = 1.0 / 1.0;
    PlasmaVals v;
    v.psi_inv = 1.0 / p.psi[ip];
    v.vx_c = p.ux[ip] * v.psi_inv;
    v.vy_c = p.uy[ip] * v.psi_inv;
    v.q_invvol = charge_invvol * p.w[ip];
    v.xmid = (p.x[ip] - g.xoff) * g.dxi;
    v.ymid = (p.y[ip] - g.yoff) * g.dyi;
    const double Aabssqp = 0.0;
    v.gamma_psi = 0.5 * ((1.0 + 0.5 * Aabssqp) * v.psi_inv * v.psi_inv + v.vx_c * v.vx_c * clightinv * clightinv +
                         v.vy_c * v.vy_c * clightinv * clightinv + 1.0);
    return v;
}

// The beam deposit's per-particle values (BeamDepositCurrent.cpp:133-193): [0] wq vx,
// [1] wq vy, [2] wq vz, then the shape factors of its O + 1 stencil columns and of its O + 1
// stencil rows (a beam's dense core makes every cell merge long lists: the factors are taken
// once per particle here instead of once per cell there).
constexpr int BEAM_NREC = 3 + 2 * 4;

// slot of a particle in the list of start cell k (k < 0: none): one atomic per distinct k
// among the active lanes of the wavefront, and consecutive slots for its lanes in lane order
__device__ __forceinline__ int list_slot(int k, int32_t* __restrict__ count) {
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    unsigned long long todo = __ballot(k >= 0);
    unsigned long long peers = 0;
    while (todo) {
        const int l = __ffsll((long long)todo) - 1;
        const int kk = __shfl(k, l);
        const unsigned long long m = __ballot(k == kk);
        if (k == kk) peers = m;
        todo &= ~m;
    }
    const int leader = peers ? __ffsll((long long)peers) - 1 : lane;
    int base = 0;
    if (k >= 0 && lane == leader) base = atomicAdd(&count[k], __popcll(peers));
    base = __shfl(base, leader);
    return base + __popcll(peers & lt);
}

// Plasma keys of the current deposition: the start cell of every valid particle, and the QSA
// removal of the deposit functor (PlasmaDepositCurrent.cpp:182-198).
template <int O>
__global__ __launch_bounds__(256) void k_keys_plasma(Geom g, hpq_plasma p, CLView cl, double charge_invvol,
                                                     double max_qsa, int check_qsa, int32_t* __restrict__ key,
                                                     int32_t* __restrict__ slot, int32_t* __restrict__ count,
                                                     int* __restrict__ qsa_count) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= p.n) return;
    int k = -1;
    if (p.valid[ip]) {
        const PlasmaVals v = plasma_vals(g, p, ip, charge_invvol);
        if (check_qsa && (v.gamma_psi < 0.0 || v.gamma_psi > max_qsa || v.psi_inv < 0.0)) {
            atomicAdd(qsa_count, 1);
            p.w[ip] = 0.0;
            p.valid[ip] = 0;
        } else {
            const int bx = shape<O>(v.xmid, 0).cell - shape_delta<O>() - cl.bxlo;
            const int by = shape<O>(v.ymid, 0).cell - shape_delta<O>() - cl.bylo;
            if (bx >= 0 && bx < cl.wx && by >= 0 && by < cl.wy) k = by * cl.wx + bx;
        }
    }
    key[ip] = k;
    const int sl = list_slot(k, count);
    if (k >= 0) slot[ip] = sl;
}

// Plasma keys of the explicit deposition with a start cell of its own
// (ExplicitDeposition.cpp:109-141).
template <int T, int O>
__global__ __launch_bounds__(256) void k_keys_explicit(Geom g, hpq_plasma p, CLView cl, int32_t* __restrict__ key,
                                                       int32_t* __restrict__ slot, int32_t* __restrict__ count) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= p.n) return;
    int k = -1;
    if (p.valid[ip]) {
        const double xmid = (p.x[ip] - g.xoff) * g.dxi;
        const double ymid = (p.y[ip] - g.yoff) * g.dyi;
        const int bx = dshape<T, O>(xmid, 0).cell - dshape_delta<T, O>() - cl.bxlo;
        const int by = dshape<T, O>(ymid, 0).cell - dshape_delta<T, O>() - cl.bylo;
        if (bx >= 0 && bx < cl.wx && by >= 0 && by < cl.wy) k = by * cl.wx + bx;
    }
    key[ip] = k;
    const int sl = list_slot(k, count);
    if (k >= 0) slot[ip] = sl;
}

// Beam keys (BeamDepositCurrent.cpp:109-160): particles first .. first + count - 1 of the
// beam, numbered from 0.
template <int O>
__global__ __launch_bounds__(256) void k_keys_beam(Geom g, hpq_beam b, long first, long count_p, CLView cl,
                                                   const int32_t* __restrict__ select, int select_value,
                                                   int32_t* __restrict__ key, int32_t* __restrict__ slot,
                                                   int32_t* __restrict__ count) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count_p) return;
    const long ip = first + t;
    int k = -1;
    if (b.valid[ip] && !(select && select[ip] != select_value)) {
        const double xmid = (b.x[ip] - g.xoff) * g.dxi;
        const double ymid = (b.y[ip] - g.yoff) * g.dyi;
        const int bx = shape<O>(xmid, 0).cell - shape_delta<O>() - cl.bxlo;
        const int by = shape<O>(ymid, 0).cell - shape_delta<O>() - cl.bylo;
        if (bx >= 0 && bx < cl.wx && by >= 0 && by < cl.wy) k = by * cl.wx + bx;
    }
    key[t] = k;
    const int sl = list_slot(k, count);
    if (k >= 0) slot[t] = sl;
}

__global__ __launch_bounds__(256) void k_zero_int(int32_t* __restrict__ a, long n) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) a[t] = 0;
}

// the two overflow counts (ovf[0], ovf[nb]) of a fresh set of lists
__global__ void k_zero_ovf(int32_t* __restrict__ ovf, long nb) {
    if (threadIdx.x == 0) {
        ovf[0] = 0;
        ovf[nb] = 0;
    }
}

// particle t into its list at its atomic slot
__global__ __launch_bounds__(256) void k_lists_scatter(long n, const int32_t* __restrict__ key,
                                                       const int32_t* __restrict__ slot,
                                                       const int32_t* __restrict__ off, int32_t* __restrict__ unsorted) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n || key[t] < 0) return;
    unsorted[off[key[t]] + slot[t]] = int32_t(t);
}

// the rank of particle t by index inside its list: its list position
__device__ __forceinline__ int list_position(long t, const int32_t* __restrict__ key, const int32_t* __restrict__ off,
                                             const int32_t* __restrict__ unsorted) {
    const int a = off[key[t]], b = off[key[t] + 1];
    int r = 0;
    for (int u = a; u < b; ++u) r += unsorted[u] < int32_t(t) ? 1 : 0;
    return a + r;
}

// plasma particle t to its list position, with its record (rec: PLASMA_NREC arrays of
// n_rec doubles, list order)
__global__ __launch_bounds__(256) void k_lists_order_plasma(Geom g, hpq_plasma p, double charge_invvol,
                                                            const int32_t* __restrict__ key,
                                                            const int32_t* __restrict__ off,
                                                            const int32_t* __restrict__ unsorted,
                                                            int32_t* __restrict__ perm, double* __restrict__ rec,
                                                            long n_rec) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= p.n || key[t] < 0) return;
    const int pos = list_position(t, key, off, unsorted);
    perm[pos] = int32_t(t);
    const PlasmaVals v = plasma_vals(g, p, t, charge_invvol);
    rec[pos] = v.q_invvol;
    rec[n_rec + pos] = v.vx_c;
    rec[2 * n_rec + pos] = v.vy_c;
    rec[3 * n_rec + pos] = v.gamma_psi;
    rec[4 * n_rec + pos] = v.psi_inv;
    rec[5 * n_rec + pos] = v.xmid;
    rec[6 * n_rec + pos] = v.ymid;
}

// beam particle first + t to its list position, with its record (BEAM_NREC arrays, list order)
template <int O>
__global__ __launch_bounds__(256) void k_lists_order_beam(Geom g, hpq_beam b, long first, long count_p,
                                                          const int32_t* __restrict__ key,
                                                          const int32_t* __restrict__ off,
                                                          const int32_t* __restrict__ unsorted,
                                                          int32_t* __restrict__ perm, double* __restrict__ rec) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count_p || key[t] < 0) return;
    const int pos = list_position(t, key, off, unsorted);
    perm[pos] = int32_t(t);
    const long ip = first + t;
    const double clightsq = 1.0 / (1.0 * 1.0);
    const double ux = b.ux[ip], uy = b.uy[ip], uz = b.uz[ip];
    const double gaminv = 1.0 / sqrt(1.0 + ux * ux * clightsq + uy * uy * clightsq + uz * uz * clightsq);
    const double wq = b.charge * b.w[ip] * 1.0;
    rec[pos] = wq * (ux * gaminv);
    rec[count_p + pos] = wq * (uy * gaminv);
    rec[2 * count_p + pos] = wq * (uz * gaminv);
    const double xmid = (b.x[ip] - g.xoff) * g.dxi;
    const double ymid = (b.y[ip] - g.yoff) * g.dyi;
#pragma unroll
    for (int k = 0; k <= O; ++k) {
        rec[(3 + k) * count_p + pos] = shape<O>(xmid, k).s;
        rec[(3 + O + 1 + k) * count_p + pos] = shape<O>(ymid, k).s;
    }
}

// ---- gathers ------------------------------------------------------------------------------

struct PlasmaRec {
    double q_invvol, vx_c, vy_c, gamma_psi, psi_inv, xmid, ymid;
};
__device__ __forceinline__ PlasmaRec load_plasma_rec(const double* __restrict__ rec, long n, int pos) {
    return {rec[pos], rec[n + pos], rec[2 * n + pos], rec[3 * n + pos], rec[4 * n + pos], rec[5 * n + pos],
            rec[6 * n + pos]};
}
struct BeamRec {
    double wqx, wqy, wqz, sx, sy;
};

// ::DepositCurrent's per-cell sums, PlasmaDepositCurrent.cpp:199-283: comps jx, jy, jz, rho,
// chi, rhomjz (-1: skip; jx and jy go together)
template <int O, bool STAGED>
__global__ __launch_bounds__(GS * GS) void k_gather_current(Geom g, CLView cl, const double* __restrict__ rec, long n,
                                                            double* __restrict__ store, int cjx, int cjy, int cjz,
                                                            int crho, int cchi, int crm,
                                                            double charge_mu0_mass_ratio, int32_t* __restrict__ ovf) {
    constexpr int S = O + 1, DELTA = shape_delta<O>(), CAP = GS_CAP;
    __shared__ StagedLists<CAP> sl;
    int b, i0, j0;
    if (!gather_setup<STAGED, S, DELTA>(g, cl, sl, ovf, b, i0, j0)) return;
    const int li = int(threadIdx.x % GS), lj = int(threadIdx.x / GS);
    const int i = i0 + li, j = j0 + lj;
    const bool in = i <= g.nx - 1 + g.ng && j <= g.ny - 1 + g.ng;
    if (STAGED && !block_raster_ordered<S>(g, cl, sl, in, i, j, li, lj, b, ovf)) return;
    if (!in) return;
    const long k = g.idx(i, j);
    const long ps = g.plane;
    double ajx = cjx >= 0 ? store[cjx * ps + k] : 0.0, ajy = cjy >= 0 ? store[cjy * ps + k] : 0.0;
    double ajz = cjz >= 0 ? store[cjz * ps + k] : 0.0, arho = crho >= 0 ? store[crho * ps + k] : 0.0;
    double achi = cchi >= 0 ? store[cchi * ps + k] : 0.0, arm = crm >= 0 ? store[crm * ps + k] : 0.0;
    int hits = 0;
    const double clight = 1.0;
    const double q_mu0_mass_ratio = charge_mu0_mass_ratio;
    walk_cell<S, DELTA, STAGED>(g, cl, sl, i, j, li, lj, [&](int pos, int, int) { return load_plasma_rec(rec, n, pos); },
                                  [&](const PlasmaRec& r, int ix, int iy) {
                                      const double sx = shape<O>(r.xmid, ix).s;
                                      const double sy = shape<O>(r.ymid, iy).s;
                                      const double charge_density = r.q_invvol * sx * sy;
                                      if (cjx >= 0) {
                                          ajx += charge_density * r.vx_c;
                                          ajy += charge_density * r.vy_c;
                                      }
                                      if (cjz >= 0) ajz += charge_density * (r.gamma_psi - 1.0) * clight;
                                      if (crho >= 0) arho += charge_density * r.gamma_psi;
                                      if (cchi >= 0) achi += charge_density * q_mu0_mass_ratio * r.psi_inv;
                                      if (crm >= 0) arm += charge_density;
                                      ++hits;
                                  });
    if (hits == 0) return;
    if (cjx >= 0) {
        store[cjx * ps + k] = ajx;
        store[cjy * ps + k] = ajy;
    }
    if (cjz >= 0) store[cjz * ps + k] = ajz;
    if (crho >= 0) store[crho * ps + k] = arho;
    if (cchi >= 0) store[cchi * ps + k] = achi;
    if (crm >= 0) store[crm * ps + k] = arm;
}

// ::ExplicitDeposition's per-cell sums, ExplicitDeposition.cpp:186-262 (no laser), with the
// cell's own Bz, Ez, ExmBy, EypBx
template <int T, int O, bool STAGED>
__global__ __launch_bounds__(GS * GS) void k_gather_explicit(Geom g, CLView cl, const double* __restrict__ rec, long n,
                                                             double* __restrict__ store, int cBz, int cEz, int cExmBy,
                                                             int cEypBx, int cSy, int cSx, double charge_mass_ratio,
                                                             int32_t* __restrict__ ovf) {
    constexpr int S = dshape_width<T, O>(), DELTA = dshape_delta<T, O>(), CAP = GS_CAP;
    __shared__ StagedLists<CAP> sl;
    int b, i0, j0;
    if (!gather_setup<STAGED, S, DELTA>(g, cl, sl, ovf, b, i0, j0)) return;
    const int li = int(threadIdx.x % GS), lj = int(threadIdx.x / GS);
    const int i = i0 + li, j = j0 + lj;
    const bool in = i <= g.nx - 1 + g.ng && j <= g.ny - 1 + g.ng;
    if (STAGED && !block_raster_ordered<S>(g, cl, sl, in, i, j, li, lj, b, ovf)) return;
    if (!in) return;
    const long k = g.idx(i, j);
    const long ps = g.plane;
    const double Bz_v = store[cBz * ps + k], Ez_v = store[cEz * ps + k];
    const double ExmBy_v = store[cExmBy * ps + k], EypBx_v = store[cEypBx * ps + k];
    double aSy = store[cSy * ps + k], aSx = store[cSx * ps + k];
    int hits = 0;
    const double clight = 1.0, clight_inv = 1.0 / 1.0;
    const double dxi = g.dxi, dyi = g.dyi;
    const double q_mass_ratio = charge_mass_ratio;
    walk_cell<S, DELTA, STAGED>(
        g, cl, sl, i, j, li, lj, [&](int pos, int, int) { return load_plasma_rec(rec, n, pos); },
        [&](const PlasmaRec& r, int ix, int iy) {
            if (T == 2 && (ix == 0 || ix == O + 2) && (iy == 0 || iy == O + 2)) return;
            const double charge_density_mu0 = r.q_invvol, vx = r.vx_c, vy = r.vy_c;
            const double gamma_psi = r.gamma_psi, psi_inv = r.psi_inv;
            const Dsf sy = dshape<T, O>(r.ymid, iy);
            const Dsf sx = dshape<T, O>(r.xmid, ix);
            const double AabssqDxp = 0.0, AabssqDyp = 0.0;
            aSy += charge_density_mu0 * (
                -sx.s * sy.s * (
                    -Bz_v * vx
                    + (Ez_v * vy
                    + ExmBy_v * (-vx * vy)
                    + EypBx_v * (gamma_psi - vy * vy)) * clight_inv
                    - 0.25 * AabssqDyp * q_mass_ratio * psi_inv
                ) * q_mass_ratio * psi_inv
                + (-sx.ds * sy.s * dxi * (-vx * vy)
                   - sx.s * sy.ds * dyi * (gamma_psi - vy * vy - 1.0)) * clight);
            aSx += charge_density_mu0 * (
                +sx.s * sy.s * (
                    +Bz_v * vy
                    + (Ez_v * vx
                    + ExmBy_v * (gamma_psi - vx * vx)
                    + EypBx_v * (-vx * vy)) * clight_inv
                    - 0.25 * AabssqDxp * q_mass_ratio * psi_inv
                ) * q_mass_ratio * psi_inv
                + (+sx.ds * sy.s * dxi * (gamma_psi - vx * vx - 1.0)
                   + sx.s * sy.ds * dyi * (-vx * vy)) * clight);
            ++hits;
        });
    if (hits == 0) return;
    store[cSy * ps + k] = aSy;
    store[cSx * ps + k] = aSx;
}

// ::DepositCurrentSlice's per-cell sums, BeamDepositCurrent.cpp:160-193
template <int O, bool STAGED>
__global__ __launch_bounds__(GS * GS) void k_gather_beam(Geom g, CLView cl, const double* __restrict__ rec, long n,
                                                         double* __restrict__ store, int cjx, int cjy, int cjz,
                                                         int32_t* __restrict__ ovf) {
    constexpr int S = O + 1, DELTA = shape_delta<O>(), CAP = GS_CAP_BEAM;
    __shared__ StagedLists<CAP> sl;
    int b, i0, j0;
    if (!gather_setup<STAGED, S, DELTA>(g, cl, sl, ovf, b, i0, j0)) return;
    const int li = int(threadIdx.x % GS), lj = int(threadIdx.x / GS);
    const int i = i0 + li, j = j0 + lj;
    const bool in = i <= g.nx - 1 + g.ng && j <= g.ny - 1 + g.ng;
    if (STAGED && !block_raster_ordered<S>(g, cl, sl, in, i, j, li, lj, b, ovf)) return;
    if (!in) return;
    const long k = g.idx(i, j);
    const long ps = g.plane;
    double ajx = cjx >= 0 ? store[cjx * ps + k] : 0.0, ajy = cjy >= 0 ? store[cjy * ps + k] : 0.0;
    double ajz = cjz >= 0 ? store[cjz * ps + k] : 0.0;
    int hits = 0;
    walk_cell<S, DELTA, STAGED>(
        g, cl, sl, i, j, li, lj,
        [&](int pos, int ix, int iy) {
            return BeamRec{rec[pos], rec[n + pos], rec[2 * n + pos], rec[(3 + ix) * n + pos],
                           rec[(3 + O + 1 + iy) * n + pos]};
        },
        [&](const BeamRec& r, int, int) {
            if (cjx >= 0) {
                ajx += r.sx * r.sy * r.wqx;
                ajy += r.sx * r.sy * r.wqy;
            }
            if (cjz >= 0) ajz += r.sx * r.sy * r.wqz;
            ++hits;
        });
    if (hits == 0) return;
    if (cjx >= 0) {
        store[cjx * ps + k] = ajx;
        store[cjy * ps + k] = ajy;
    }
    if (cjz >= 0) store[cjz * ps + k] = ajz;
}

template <class T>
void grow(T*& p, long& cap, long need) {
    if (need <= cap) return;
    if (p) HPQ_HIP(hipFree(p));
    p = nullptr;
    HPQ_HIP(hipMalloc(&p, sizeof(T) * size_t(std::max(need, 1L))));
    cap = need;
}

// base-cell grid of a context: every start cell whose stencil (at most 6 wide, start at most 2
// below the base cell) reaches the guarded plane; the deposition tiles
void lists_geometry(const hpq_ctx* c, CellLists& L) {
    const Geom& g = c->g;
    L.bxlo = -g.ng - 6;
    L.bylo = -g.ng - 6;
    L.wx = g.nx + 2 * g.ng + 9;
    L.wy = g.ny + 2 * g.ng + 9;
    L.T = c->p.tile_size > 0 ? c->p.tile_size : 32;
    L.ntx = (g.ldx + L.T - 1) / L.T;
    L.nty = (g.ldy + L.T - 1) / L.T;
}

CLView view(const CellLists& L) {
    return CLView{L.bxlo, L.bylo, L.wx, L.wy, L.T, L.ntx, L.nty, L.off, L.perm};
}

// gather blocks over the guarded plane
long gather_blocks(const Geom& g) { return long((g.ldx + GS - 1) / GS) * ((g.ldy + GS - 1) / GS); }

void lists_reserve(const Geom& g, CellLists& L, long n, int nrec) {
    const long ns = long(L.wx) * L.wy;
    // overflow lists of the gathers, [count, blocks...] each: the current (or beam)
    // deposition's at ovf, the explicit deposition's at ovf + gather_blocks + 1
    grow(L.ovf, L.ovf_cap, 2 * (gather_blocks(g) + 1));
    grow(L.off, L.off_cap, ns + 1);
    grow(L.tmp, L.tmp_cap, scan_tmp_size(ns + 1));
    grow(L.key, L.key_cap, n);
    grow(L.slot, L.slot_cap, n);
    grow(L.unsorted, L.unsorted_cap, n);
    grow(L.perm, L.perm_cap, n);
    grow(L.rec, L.rec_cap, long(nrec) * n);
}

// count (already in L.off) -> offsets -> unsorted lists; the caller's order kernel then
// sorts every list and writes the records
void lists_scatter(hipStream_t s, CellLists& L, long n) {
    const long ns = long(L.wx) * L.wy;
    scan_exclusive(s, L.off, ns + 1, L.tmp);
    hipLaunchKernelGGL(k_lists_scatter, grid1d(n, 256), dim3(256), 0, s, n, L.key, L.slot, L.off, L.unsorted);
    HPQ_HIP(hipGetLastError());
}

void lists_clear(const Geom& g, hipStream_t s, CellLists& L) {
    const long ns = long(L.wx) * L.wy;
    hipLaunchKernelGGL(k_zero_int, grid1d(ns + 1, 256), dim3(256), 0, s, L.off, ns + 1);
    hipLaunchKernelGGL(k_zero_ovf, dim3(1), dim3(64), 0, s, L.ovf, gather_blocks(g) + 1);
}

dim3 gather_grid(const Geom& g) { return dim3(unsigned(gather_blocks(g))); }

int shape_base_of(int order) {
    switch (order) {
        case 0: return shape_base<0>();
        case 1: return shape_base<1>();
        case 2: return shape_base<2>();
        default: return shape_base<3>();
    }
}

}  // namespace

void cell_lists_free(CellLists& L) {
    for (void* q : {(void*)L.off, (void*)L.tmp, (void*)L.key, (void*)L.slot, (void*)L.unsorted, (void*)L.perm,
                    (void*)L.rec, (void*)L.ovf})
        if (q) (void)hipFree(q);
    L = CellLists{};
}

// ::DepositCurrent in the CPU build's order; builds the plasma lists (L.state names the
// particle state they describe, for the explicit deposition that follows on the same state)
void deposit_current_ordered(hpq_ctx* c, const hpq_plasma* p, double* store, const int32_t* comps, bool ions) {
    const Geom& g = c->g;
    CellLists& L = c->cl_plasma;
    L.valid = false;
    if (p->n <= 0) return;
    if (p->n > INT32_MAX) throw HpqError(HPQ_E_INVALID, "ordered deposition: more than 2^31-1 particles");
    lists_geometry(c, L);
    lists_reserve(g, L, p->n, PLASMA_NREC);
    const double charge = ions ? -c->p.plasma_charge : c->p.plasma_charge;
    const double mu0 = 1.0;
    const double invvol = g.dx * g.dy * g.dxi * g.dyi;
    const double charge_invvol = charge * invvol;
    const double cmr = charge * mu0 / c->p.plasma_mass;
    TraceScope ts(c, HPQ_TRACE_DEPOSIT);
    lists_clear(g, c->stream, L);
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        hipLaunchKernelGGL(k_keys_plasma<o>, grid1d(p->n, 256), dim3(256), 0, c->stream, g, *p, view(L), charge_invvol,
                           c->p.max_qsa, 1, L.key, L.slot, L.off, c->d_count);
        HPQ_HIP(hipGetLastError());
        lists_scatter(c->stream, L, p->n);
        hipLaunchKernelGGL(k_lists_order_plasma, grid1d(p->n, 256), dim3(256), 0, c->stream, g, *p, charge_invvol,
                           L.key, L.off, L.unsorted, L.perm, L.rec, p->n);
        // the fast pass, then the blocks it listed (merged)
        hipLaunchKernelGGL((k_gather_current<o, true>), gather_grid(g), dim3(GS * GS), 0, c->stream, g, view(L),
                           L.rec, p->n, store, comps[0], comps[1], comps[2], comps[3], comps[4], comps[5], cmr, L.ovf);
        hipLaunchKernelGGL((k_gather_current<o, false>), gather_grid(g), dim3(GS * GS), 0, c->stream, g, view(L),
                           L.rec, p->n, store, comps[0], comps[1], comps[2], comps[3], comps[4], comps[5], cmr, L.ovf);
    });
    HPQ_HIP(hipGetLastError());
    L.valid = !ions;
    L.base = shape_base_of(c->p.depos_order);
    L.x = p->x;
    L.n = p->n;
}

// ::ExplicitDeposition in the CPU build's order; reuses the lists of the current deposition
// on the same particle state when the start cells share their floor
void explicit_deposition_ordered(hpq_ctx* c, const hpq_plasma* p, double* store) {
    const Geom& g = c->g;
    const Layout& Ly = c->L;
    CellLists& L = c->cl_plasma;
    if (p->n <= 0) return;
    if (p->n > INT32_MAX) throw HpqError(HPQ_E_INVALID, "ordered deposition: more than 2^31-1 particles");
    const double mu0 = 1.0;
    const double invvol = g.dx * g.dy * g.dxi * g.dyi;
    const double qimu0 = c->p.plasma_charge * invvol * mu0;
    const double qmr = c->p.plasma_charge / c->p.plasma_mass;
    const int T = c->p.derivative_type;
    TraceScope ts(c, HPQ_TRACE_EXPLICIT_DEPOSIT);
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        auto run = [&](auto TT) {
            constexpr int t = decltype(TT)::value;
            const bool reuse = L.valid && L.x == p->x && L.n == p->n && L.base == dshape_base<t, o>();
            if (!reuse) {
                lists_geometry(c, L);
                lists_reserve(g, L, p->n, PLASMA_NREC);
                lists_clear(g, c->stream, L);
                hipLaunchKernelGGL((k_keys_explicit<t, o>), grid1d(p->n, 256), dim3(256), 0, c->stream, g, *p, view(L),
                                   L.key, L.slot, L.off);
                HPQ_HIP(hipGetLastError());
                lists_scatter(c->stream, L, p->n);
                hipLaunchKernelGGL(k_lists_order_plasma, grid1d(p->n, 256), dim3(256), 0, c->stream, g, *p, qimu0,
                                   L.key, L.off, L.unsorted, L.perm, L.rec, p->n);
                L.valid = true;
                L.base = dshape_base<t, o>();
                L.x = p->x;
                L.n = p->n;
            }
            // the explicit deposition's overflow list (its count cleared here: the lists may be
            // reused by several explicit depositions)
            int32_t* ovf = L.ovf + gather_blocks(g) + 1;
            hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, ovf, 1L);
            hipLaunchKernelGGL((k_gather_explicit<t, o, true>), gather_grid(g), dim3(GS * GS), 0, c->stream, g,
                               view(L), L.rec, p->n, store, Ly.Bz, Ly.Ez, Ly.ExmBy, Ly.EypBx, Ly.Sy, Ly.Sx, qmr, ovf);
            hipLaunchKernelGGL((k_gather_explicit<t, o, false>), gather_grid(g), dim3(GS * GS), 0, c->stream, g,
                               view(L), L.rec, p->n, store, Ly.Bz, Ly.Ez, Ly.ExmBy, Ly.EypBx, Ly.Sy, Ly.Sx, qmr, ovf);
        };
        if (T == 0) run(std::integral_constant<int, 0>{});
        else if (T == 1) run(std::integral_constant<int, 1>{});
        else run(std::integral_constant<int, 2>{});
    });
    HPQ_HIP(hipGetLastError());
}

// ::DepositCurrentSlice in the CPU build's order (particles of `slice` whose select entry
// equals select_value; select NULL: all)
void deposit_beam_ordered(hpq_ctx* c, const hpq_beam* b, int slice, double* store, const int32_t* comps,
                          const int32_t* select, int select_value) {
    const Geom& g = c->g;
    CellLists& L = c->cl_beam;
    const long first = b->off[slice];
    const long count = b->off[slice + 1] - first;
    if (count <= 0) return;
    if (count > INT32_MAX) throw HpqError(HPQ_E_INVALID, "ordered deposition: more than 2^31-1 particles");
    lists_geometry(c, L);
    lists_reserve(g, L, count, BEAM_NREC);
    TraceScope ts(c, HPQ_TRACE_BEAM_DEPOSIT);
    lists_clear(g, c->stream, L);
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        hipLaunchKernelGGL(k_keys_beam<o>, grid1d(count, 256), dim3(256), 0, c->stream, g, *b, first, count, view(L),
                           select, select_value, L.key, L.slot, L.off);
        HPQ_HIP(hipGetLastError());
        lists_scatter(c->stream, L, count);
        hipLaunchKernelGGL(k_lists_order_beam<o>, grid1d(count, 256), dim3(256), 0, c->stream, g, *b, first, count,
                           L.key, L.off, L.unsorted, L.perm, L.rec);
        hipLaunchKernelGGL((k_gather_beam<o, true>), gather_grid(g), dim3(GS * GS), 0, c->stream, g, view(L),
                           L.rec, count, store, comps[0], comps[1], comps[2], L.ovf);
        hipLaunchKernelGGL((k_gather_beam<o, false>), gather_grid(g), dim3(GS * GS), 0, c->stream, g, view(L),
                           L.rec, count, store, comps[0], comps[1], comps[2], L.ovf);
    });
    HPQ_HIP(hipGetLastError());
}

}  // namespace hpq
