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
// particle's record position, (ix, iy) its stencil offset that lands on (i, j). The lists of
// the mask SKIP (bit ky * S + kx), whose contributions the caller drops, are left out. A list's
// end is read from the source when the list is taken, not kept: the merge's registers bound
// how many wavefronts share a SIMD.
template <int S, unsigned long long SKIP = 0ull, class Src, class F>
__device__ __forceinline__ void walk_ordered(const Geom& g, const CLView& cl, int i, int j, const Src& src, F&& f) {
    const int x0 = i - S + 1, y0 = j - S + 1;
    int h[S * S];
    bool ne[S * S];   // list non-empty
#pragma unroll
    for (int ky = 0; ky < S; ++ky) {
        int o[S + 1];
#pragma unroll
        for (int kx = 0; kx <= S; ++kx) o[kx] = src.bound(kx, ky);
#pragma unroll
        for (int kx = 0; kx < S; ++kx) {
            h[ky * S + kx] = o[kx];
            ne[ky * S + kx] = o[kx] < o[kx + 1];
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
            const bool act = !((SKIP >> k) & 1ull) && sx >= ax && sx <= bx && sy >= ay && sy <= by && ne[k];
            v[k] = act ? src.idx(h[k]) : INT_MAX;
        }
        while (true) {
            int m1 = INT_MAX, m2 = INT_MAX, km = -1;
#pragma unroll
            for (int k = 0; k < S * S; ++k) {
                if ((SKIP >> k) & 1ull) continue;
                if (v[k] < m1) {
                    m2 = m1;
                    m1 = v[k];
                    km = k;
                } else if (v[k] < m2) {
                    m2 = v[k];
                }
            }
            if (km < 0) break;
            int hc = 0;
#pragma unroll
            for (int k = 0; k < S * S; ++k)
                if (!((SKIP >> k) & 1ull) && k == km) hc = h[k];
            const int kyl = km / S, kxl = km - kyl * S;
            const int ec = src.bound(kxl + 1, kyl);
            const int ix = S - 1 - kxl, iy = S - 1 - kyl;
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
                if (!((SKIP >> k) & 1ull) && k == km) { h[k] = hc; v[k] = p; }
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
constexpr int GS_CAP_REC = 2560;   // of the plasma fast passes, with their records (150 KiB)
template <int CAP, bool IDX = true>
struct StagedLists {
    int off[GS_MAXW][GS_MAXW + 1];   // window list bounds, in staged entries
    int base[GS_MAXW];               // record position = staged entry + base[row]
    int total;                       // staged entries (more than CAP: not staged)
    int idx[IDX ? CAP : 1];          // particle index of every staged entry (IDX)
};
// staged lists with the NF record fields of every staged entry (NF = 0: indices only)
template <int CAP, int NF>
struct StagedRecs {
    StagedLists<CAP> l;
    double rec[NF][CAP];
};
template <int CAP>
struct StagedRecs<CAP, 0> {
    StagedLists<CAP> l;
};
template <int CAP>
struct StagedView {
    const StagedLists<CAP>& s;
    int li, lj;   // the cell's column and row in the block
    __device__ __forceinline__ int bound(int kx, int ky) const { return s.off[lj + ky][li + kx]; }
    __device__ __forceinline__ int idx(int h) const { return s.idx[h]; }
    __device__ __forceinline__ int pos(int h, int ky) const { return h + s.base[lj + ky]; }
};

// blocks of GS x BH cells: BH = GS for the merge passes and the per-list walks, FLAT_BH for the
// flat-walk fast passes (two of their workgroups fit in a CU's LDS)
constexpr int FLAT_BH = 8;
__device__ __forceinline__ int gather_nbx(const Geom& g) { return (g.ldx + GS - 1) / GS; }
template <int BH>
__device__ __forceinline__ int gather_nby(const Geom& g) { return (g.ldy + BH - 1) / BH; }

// stages the rows of lists below block (i0, j0) of GS x BH cells for stencils of S cells
// starting DELTA from the base cell, and calls copy(t, pos) for every staged entry t (record
// position pos); every thread of the workgroup calls it
template <int S, int DELTA, int BH, int CAP, bool IDX, class Copy>
__device__ __forceinline__ void stage_lists(const CLView& cl, int i0, int j0, StagedLists<CAP, IDX>& sl, Copy&& copy) {
    constexpr int NW = GS + S - 1, NH = BH + S - 1;
    static_assert(NW <= GS_MAXW && NH <= GS_MAXW, "window wider than the staging area");
    const int col0 = i0 - S + 1 - DELTA - cl.bxlo;
    const int row0 = j0 - S + 1 - DELTA - cl.bylo;
    // list bounds; lists outside the list grid are empty
    for (int e = threadIdx.x; e < NH * (NW + 1); e += blockDim.x) {
        const int r = e / (NW + 1), c = e - r * (NW + 1);
        const int row = row0 + r;
        sl.off[r][c] = (row >= 0 && row < cl.wy) ? cl.off[row * cl.wx + min(max(col0 + c, 0), cl.wx)] : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int r = 0; r < NH; ++r) {
            const int a = sl.off[r][0];
            sl.base[r] = a - run;
            run += sl.off[r][NW] - a;
        }
        sl.total = run;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NH * (NW + 1); e += blockDim.x) {
        const int r = e / (NW + 1), c = e - r * (NW + 1);
        sl.off[r][c] -= sl.base[r];
    }
    __syncthreads();
    const int total = sl.total;
    if (total <= CAP) {
        // U entries per thread at a time: their rows first, then all their loads in flight
        constexpr int U = 4;
        for (int t0 = threadIdx.x; t0 < total; t0 += U * blockDim.x) {
            int pos[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = t0 + u * int(blockDim.x);
                // the row holding staged entry t: the last row starting at or before t
                int lo = 0, hi = NH - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (sl.off[mid][0] <= t) lo = mid;
                    else hi = mid - 1;
                }
                pos[u] = t < total ? t + sl.base[lo] : -1;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = t0 + u * int(blockDim.x);
                if (pos[u] >= 0) {
                    if constexpr (IDX) sl.idx[t] = cl.perm[pos[u]];
                    copy(t, pos[u]);
                }
            }
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
template <bool FAST, int S, int DELTA, int BH, int CAP, bool IDX, class Copy>
__device__ __forceinline__ bool gather_setup(const Geom& g, const CLView& cl, StagedLists<CAP, IDX>& sl,
                                             int32_t* __restrict__ ovf, int& b, int& i0, int& j0, Copy&& copy) {
    const int nbx = gather_nbx(g);
    if constexpr (FAST) {
        b = xcd_block(int(blockIdx.x), nbx * gather_nby<BH>(g));
    } else {
        if (int(blockIdx.x) >= ovf[0]) return false;
        b = ovf[1 + blockIdx.x];
    }
    i0 = (b % nbx) * GS - g.ng;
    j0 = (b / nbx) * BH - g.ng;
    stage_lists<S, DELTA, BH>(cl, i0, j0, sl, copy);
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

// The walk of cell (i, j) = block cell (li, lj) in the CPU build's order, add(load(at), ix, iy)
// for every contribution. FAST: the lists in raster order (stencil offsets known at compile
// time, records read ahead by unrolling); `at` is the staged entry when ENTRY (records staged
// in LDS), the record position otherwise. MERGE: the merge of the lists, from LDS when the
// block's rows were staged, from global memory otherwise; `at` is the record position.
template <int S, int DELTA, bool FAST, bool ENTRY, unsigned long long SKIP = 0ull, int CAP, class Load, class Add>
__device__ __forceinline__ void walk_cell(const Geom& g, const CLView& cl, const StagedLists<CAP>& sl, int i, int j,
                                          int li, int lj, Load&& load, Add&& add) {
    using Rec = decltype(load(0, 0, 0));
    const StagedView<CAP> src{sl, li, lj};
    if constexpr (FAST) {
        // colours at run time (most windows hold one), lists unrolled: one copy of the
        // contribution code per stencil offset
        const WindowColours<S> wc(g, cl, i, j);
#pragma unroll 1
        for (int c = 0; c < 4; ++c) {
            const unsigned long long mk = c == 0 ? wc.mask[0] : c == 1 ? wc.mask[1] : c == 2 ? wc.mask[2] : wc.mask[3];
            if (!mk) continue;
#pragma unroll
            for (int k = 0; k < S * S; ++k) {
                const int kx = k % S, ky = k / S;
                if (!((mk >> k) & 1ull) || ((SKIP >> k) & 1ull)) continue;
                const int a = src.bound(kx, ky), e = src.bound(kx + 1, ky);
#pragma unroll S <= 3 ? 2 : 1
                for (int h = a; h < e; ++h)
                    add(load(ENTRY ? h : src.pos(h, ky), S - 1 - kx, S - 1 - ky), S - 1 - kx, S - 1 - ky);
            }
        }
    } else {
        auto f = pipelined<Rec>(load, add);
        if (sl.total <= CAP) {
            walk_ordered<S, SKIP>(g, cl, i, j, src, f);
        } else {
            const GlobalLists gsrc{cl, j - S + 1 - DELTA - cl.bylo, i - S + 1 - DELTA - cl.bxlo};
            walk_ordered<S, SKIP>(g, cl, i, j, gsrc, f);
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

// a fresh set of lists: the per-start-cell counts off[0, n), the two overflow counts
// (ovf[0], ovf[stride]) and, for the beam, the count of long sequences (longc[0])
__global__ __launch_bounds__(256) void k_lists_clear(int32_t* __restrict__ off, long n, int32_t* __restrict__ ovf,
                                                     long stride, int32_t* __restrict__ longc) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) off[t] = 0;
    if (t == 0) {
        ovf[0] = 0;
        ovf[stride] = 0;
        if (longc) longc[0] = 0;
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
    int u = a;
    // eight loads in flight at a time: a beam core's lists hold hundreds of particles
    for (; u + 8 <= b; u += 8) {
        int32_t q[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] = unsorted[u + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) r += q[k] < int32_t(t) ? 1 : 0;
    }
    for (; u < b; ++u) r += unsorted[u] < int32_t(t) ? 1 : 0;
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
// The staging area of a plasma gather: the fast pass stages the records with the lists (the
// block's window particles are read from global memory once, each then serves every cell of
// its stencil from LDS); the merge pass stages list indices only, for its larger blocks.
template <bool FAST>
using PlasmaStage = StagedRecs<FAST ? GS_CAP_REC : GS_CAP, FAST ? PLASMA_NREC : 0>;
template <bool FAST>
struct PlasmaRecs {
    PlasmaStage<FAST>& st;
    const double* __restrict__ rec;
    long n;
    // staged entry t <- record position pos
    __device__ __forceinline__ void operator()(int t, int pos) const {
        if constexpr (FAST) {
#pragma unroll
            for (int f = 0; f < PLASMA_NREC; ++f) st.rec[f][t] = rec[f * n + pos];
        }
    }
    // the record of staged entry `at` (FAST) or of record position `at`
    __device__ __forceinline__ PlasmaRec load(int at) const {
        if constexpr (FAST) {
            return {st.rec[0][at], st.rec[1][at], st.rec[2][at], st.rec[3][at], st.rec[4][at], st.rec[5][at],
                    st.rec[6][at]};
        } else {
            return load_plasma_rec(rec, n, at);
        }
    }
};

// threads of a plasma gather workgroup: the fast pass runs two threads per cell, each with
// its own share of the cell's sums (the cell's walk is the same for both), so each SIMD holds
// two wavefronts with the workgroup's LDS; the merge pass one thread per cell
template <bool FAST, int BH = GS>
constexpr int plasma_gather_threads() { return FAST ? 2 * GS * BH : GS * BH; }

// ::DepositCurrent's per-cell sums, PlasmaDepositCurrent.cpp:199-283: comps jx, jy, jz, rho,
// chi, rhomjz (-1: skip; jx and jy go together). Fast pass: the first thread of a cell sums
// jx, jy and jz, the second rho, chi and rhomjz.
template <int O, bool STAGED, int BH = GS>
__global__ __launch_bounds__((plasma_gather_threads<STAGED, BH>())) void k_gather_current(
    Geom g, CLView cl, const double* __restrict__ rec, long n, double* __restrict__ store, int cjx, int cjy, int cjz,
    int crho, int cchi, int crm, double charge_mu0_mass_ratio, int32_t* __restrict__ ovf) {
    constexpr int S = O + 1, DELTA = shape_delta<O>();
    __shared__ PlasmaStage<STAGED> st;
    auto& sl = st.l;
    const PlasmaRecs<STAGED> pr{st, rec, n};
    int b, i0, j0;
    if (!gather_setup<STAGED, S, DELTA, BH>(g, cl, sl, ovf, b, i0, j0, pr)) return;
    const int cell = int(threadIdx.x) % (GS * BH), half = int(threadIdx.x) / (GS * BH);
    const int li = cell % GS, lj = cell / GS;
    const int i = i0 + li, j = j0 + lj;
    const bool in = i <= g.nx - 1 + g.ng && j <= g.ny - 1 + g.ng;
    if (STAGED && !block_raster_ordered<S>(g, cl, sl, in && half == 0, i, j, li, lj, b, ovf)) return;
    if (!in) return;
    if (STAGED) {
        if (half == 0) crho = cchi = crm = -1;
        else cjx = cjy = cjz = -1;
        if (cjx < 0 && cjz < 0 && crho < 0 && cchi < 0 && crm < 0) return;
    }
    const long k = g.idx(i, j);
    const long ps = g.plane;
    double ajx = cjx >= 0 ? store[cjx * ps + k] : 0.0, ajy = cjy >= 0 ? store[cjy * ps + k] : 0.0;
    double ajz = cjz >= 0 ? store[cjz * ps + k] : 0.0, arho = crho >= 0 ? store[crho * ps + k] : 0.0;
    double achi = cchi >= 0 ? store[cchi * ps + k] : 0.0, arm = crm >= 0 ? store[crm * ps + k] : 0.0;
    int hits = 0;
    const double clight = 1.0;
    const double q_mu0_mass_ratio = charge_mu0_mass_ratio;
    walk_cell<S, DELTA, STAGED, STAGED>(g, cl, sl, i, j, li, lj, [&](int at, int, int) { return pr.load(at); },
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
// cell's own Bz, Ez, ExmBy, EypBx. Fast pass: the first thread of a cell sums Sy, the second Sx.
template <int T, int O, bool STAGED, int BH = GS>
__global__ __launch_bounds__((plasma_gather_threads<STAGED, BH>())) void k_gather_explicit(
    Geom g, CLView cl, const double* __restrict__ rec, long n, double* __restrict__ store, int cBz, int cEz, int cExmBy,
    int cEypBx, int cSy, int cSx, double charge_mass_ratio, int32_t* __restrict__ ovf) {
    constexpr int S = dshape_width<T, O>(), DELTA = dshape_delta<T, O>();
    __shared__ PlasmaStage<STAGED> st;
    auto& sl = st.l;
    const PlasmaRecs<STAGED> pr{st, rec, n};
    int b, i0, j0;
    if (!gather_setup<STAGED, S, DELTA, BH>(g, cl, sl, ovf, b, i0, j0, pr)) return;
    const int cell = int(threadIdx.x) % (GS * BH), half = int(threadIdx.x) / (GS * BH);
    const int li = cell % GS, lj = cell / GS;
    const int i = i0 + li, j = j0 + lj;
    const bool in = i <= g.nx - 1 + g.ng && j <= g.ny - 1 + g.ng;
    if (STAGED && !block_raster_ordered<S>(g, cl, sl, in && half == 0, i, j, li, lj, b, ovf)) return;
    if (!in) return;
    const bool do_sy = !STAGED || half == 0, do_sx = !STAGED || half == 1;
    const long k = g.idx(i, j);
    const long ps = g.plane;
    const double Bz_v = store[cBz * ps + k], Ez_v = store[cEz * ps + k];
    const double ExmBy_v = store[cExmBy * ps + k], EypBx_v = store[cEypBx * ps + k];
    double aSy = store[cSy * ps + k], aSx = store[cSx * ps + k];
    int hits = 0;
    const double clight = 1.0, clight_inv = 1.0 / 1.0;
    const double dxi = g.dxi, dyi = g.dyi;
    const double q_mass_ratio = charge_mass_ratio;
    // the centred derivative's corner offsets contribute nothing (see the add below): the
    // merge leaves their lists out
    constexpr unsigned long long CORNERS =
        T == 2 ? (1ull | (1ull << (S - 1)) | (1ull << (S * (S - 1))) | (1ull << (S * S - 1))) : 0ull;
    walk_cell<S, DELTA, STAGED, STAGED, CORNERS>(
        g, cl, sl, i, j, li, lj, [&](int at, int, int) { return pr.load(at); },
        [&](const PlasmaRec& r, int ix, int iy) {
            if (T == 2 && (ix == 0 || ix == O + 2) && (iy == 0 || iy == O + 2)) return;
            const double charge_density_mu0 = r.q_invvol, vx = r.vx_c, vy = r.vy_c;
            const double gamma_psi = r.gamma_psi, psi_inv = r.psi_inv;
            const Dsf sy = dshape<T, O>(r.ymid, iy);
            const Dsf sx = dshape<T, O>(r.xmid, ix);
            const double AabssqDxp = 0.0, AabssqDyp = 0.0;
            if (do_sy) {
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
            }
            if (do_sx) {
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
            }
            ++hits;
        });
    if (hits == 0) return;
    if (do_sy) store[cSy * ps + k] = aSy;
    if (do_sx) store[cSx * ps + k] = aSx;
}

// The staging area of the flat-walk fast passes: the window lists' bounds and the records of
// their entries, without the particle indices (their raster check reads those from global
// memory). Sized so that two workgroups (blocks of GS x FLAT_BH cells) fit in a CU's LDS: one
// stages its window while the other walks. A 2 x 2 lattice fills 960 of the 1400 entries of a
// centred order-2 window (20 x 12 lists).
constexpr int FLAT_CAP = 1400;                  // staged entries (77 KiB of records)
constexpr int FLAT_SLOTS = FLAT_CAP;
__device__ __forceinline__ int flat_slot(int t) { return t; }
struct FlatStage {
    StagedLists<FLAT_CAP, false> l;
    double rec[PLASMA_NREC][FLAT_SLOTS];
};
struct FlatView {
    const StagedLists<FLAT_CAP, false>& s;
    int li, lj;
    __device__ __forceinline__ int bound(int kx, int ky) const { return s.off[lj + ky][li + kx]; }
};

// fast pass of the flattened gathers: whether every cell of the block can walk its lists in
// raster order, the lists' bounds and indices read from global memory (every thread of the
// workgroup calls this; a block that cannot is listed for the merge pass)
template <int S, int DELTA>
__device__ __forceinline__ bool block_raster_ordered_global(const Geom& g, const CLView& cl, bool in, int i, int j,
                                                            int b, int32_t* __restrict__ ovf) {
    bool bad = false;
    if (in) {
        const GlobalLists src{cl, j - S + 1 - DELTA - cl.bylo, i - S + 1 - DELTA - cl.bxlo};
        bad = !raster_ordered<S>(WindowColours<S>(g, cl, i, j), src);
    }
    if (__syncthreads_or(bad)) {
        if (threadIdx.x == 0) ovf[1 + atomicAdd(&ovf[0], 1)] = b;
        return false;
    }
    return true;
}

// A cell's window lists as one sequence in the CPU build's order (colour by colour, the lists
// of a colour in raster order), the lists of the mask `skip` left out. next() moves to the
// next non-empty list: its entries [h, e) and stencil offset (ix, iy); false at the end. One
// element per loop iteration keeps the lanes of a wavefront together however their list
// lengths differ; the offset is then a run-time value.
template <int S, class View>
struct FlatWalk {
    // lists left of the current colour, and of the colours after it (shifted down as the
    // walk moves on a colour)
    unsigned long long rem, m1, m2, m3;
    View src;
    int h = 0, e = 0, ix = 0, iy = 0;
    __device__ __forceinline__ FlatWalk(const WindowColours<S>& w, const View& v, unsigned long long skip)
        : rem(w.mask[0] & ~skip), m1(w.mask[1] & ~skip), m2(w.mask[2] & ~skip), m3(w.mask[3] & ~skip), src(v) {}
    __device__ __forceinline__ bool next() {
        while (true) {
            while (rem == 0ull) {
                if ((m1 | m2 | m3) == 0ull) return false;
                rem = m1;
                m1 = m2;
                m2 = m3;
                m3 = 0ull;
            }
            const int kk = __ffsll((long long)rem) - 1;
            rem &= rem - 1ull;
            const int kx = kk % S, ky = kk / S;
            h = src.bound(kx, ky);
            e = src.bound(kx + 1, ky);
            if (h < e) {
                ix = S - 1 - kx;
                iy = S - 1 - ky;
                return true;
            }
        }
    }
};

// The order-2 shape factor (ShapeFactors.H compute_single_shape_factor, order 2) with the
// offset at run time: with t = xmid - floor(xmid + 1/2), offsets 0 and 2 are (0.5 u) u for
// u = 0.5 + sigma t, sigma = -1 and +1 (sigma t is exact and 0.5 + (-t) is 0.5 - t), offset 1
// is 0.75 - t t: the reference's expressions bit for bit.
__device__ __forceinline__ double shape2_rt(double t, double sigma, bool mid) {
    const double u = 0.5 + sigma * t;
    const double a = 0.5 * u * u;
    const double b = 0.75 - t * t;
    return mid ? b : a;
}

// The centred (type 2) order-2 derivative shape factors as one quadratic per stencil offset
// q (ShapeFactors.H single_derivative_shape_factor, centred, order 2): with
// t = (xmid + 1/2) - floor(xmid + 1/2) in [0, 1), s = (A2 t^2 + A1 t) + A0 and the reference's
// sdx = (B2 t^2 + B1 t) + B0 (ds = -sdx); row q = {A2, A1, A0, B2, B1, B0}. Evaluated in this
// form every row gives the reference's expression for its offset bit for bit: the
// coefficients 0 and +-1 multiply exactly, x + (-y) is x - y, and a vanishing term adds +0 to
// a value that is +0 or keeps its own sign. So a walk can take the offset at run time.
__constant__ double CENTRED2[5][6] = {{0.0, 0.0, 0.0, -0.25, 0.5, -0.25},
                                      {0.5, -1.0, 0.5, 0.5, -0.5, -0.25},
                                      {-1.0, 1.0, 0.5, 0.0, -0.5, 0.25},
                                      {0.5, 0.0, 0.0, -0.5, 0.5, 0.25},
                                      {0.0, 0.0, 0.0, 0.25, 0.0, 0.0}};
struct Quad {
    double a2, a1, a0, b2, b1, b0;
};

// The fast pass of ::ExplicitDeposition for the centred order-2 shape (the default
// derivative type): the same per-cell sums as k_gather_explicit<2, 2, true>, but each cell
// walks its window as one sequence (colour by colour, the lists of a colour in raster order,
// the four corner lists skipped), one element per iteration with the stencil offset at run
// time and the shape factors from the rows of CENTRED2, so the lanes of a wavefront stay
// together however their list lengths differ (the two wavefronts per SIMD cover each other's
// LDS reads). The staged records hold t of x and y in place of the cell-scaled positions.
__global__ __launch_bounds__(2 * GS * FLAT_BH) void k_gather_explicit_c2(
    Geom g, CLView cl, const double* __restrict__ rec, long n, double* __restrict__ store, int cBz, int cEz, int cExmBy,
    int cEypBx, int cSy, int cSx, double charge_mass_ratio, int32_t* __restrict__ ovf) {
    constexpr int T = 2, O = 2, S = dshape_width<T, O>(), DELTA = dshape_delta<T, O>();
    constexpr unsigned long long CORNERS = 1ull | (1ull << (S - 1)) | (1ull << (S * (S - 1))) | (1ull << (S * S - 1));
    __shared__ FlatStage st;
    __shared__ Quad quad[S];
    auto& sl = st.l;
    if (threadIdx.x < S) {
        const int q = int(threadIdx.x);
        quad[q] = Quad{CENTRED2[q][0], CENTRED2[q][1], CENTRED2[q][2], CENTRED2[q][3], CENTRED2[q][4], CENTRED2[q][5]};
    }
    auto copy = [&](int t, int pos) {
#pragma unroll
        for (int f = 0; f < 5; ++f) st.rec[f][flat_slot(t)] = rec[f * n + pos];
        const double xm = rec[5 * n + pos], ym = rec[6 * n + pos];
        const double xf = floor(xm + 0.5), yf = floor(ym + 0.5);
        st.rec[5][flat_slot(t)] = xm + 0.5 - xf;
        st.rec[6][flat_slot(t)] = ym + 0.5 - yf;
    };
    int b, i0, j0;
    if (!gather_setup<true, S, DELTA, FLAT_BH>(g, cl, sl, ovf, b, i0, j0, copy)) return;
    const int cell = int(threadIdx.x) % (GS * FLAT_BH), half = int(threadIdx.x) / (GS * FLAT_BH);
    const int li = cell % GS, lj = cell / GS;
    const int i = i0 + li, j = j0 + lj;
    const bool in = i <= g.nx - 1 + g.ng && j <= g.ny - 1 + g.ng;
    if (!block_raster_ordered_global<S, DELTA>(g, cl, in && half == 0, i, j, b, ovf)) return;
    if (!in) return;
    const long k = g.idx(i, j);
    const long ps = g.plane;
    const double Bz_v = store[cBz * ps + k], Ez_v = store[cEz * ps + k];
    const double ExmBy_v = store[cExmBy * ps + k], EypBx_v = store[cEypBx * ps + k];
    const int cout = half == 0 ? cSy : cSx;
    double acc = store[cout * ps + k];
    const double clight = 1.0, clight_inv = 1.0 / 1.0;
    const double dxi = g.dxi, dyi = g.dyi;
    const double q_mass_ratio = charge_mass_ratio;
    const double AabssqDxp = 0.0, AabssqDyp = 0.0;
    const WindowColours<S> wc(g, cl, i, j);
    const FlatView src{sl, li, lj};
    FlatWalk<S, FlatView> w(wc, src, CORNERS);
    auto load = [&](int t) {
        const int at = flat_slot(t);
        return PlasmaRec{st.rec[0][at], st.rec[1][at], st.rec[2][at], st.rec[3][at],
                         st.rec[4][at], st.rec[5][at], st.rec[6][at]};
    };
    int hits = 0;
    auto run = [&](auto HALF) {
        constexpr bool SY = decltype(HALF)::value == 0;
        Quad qx{}, qy{};
        while (true) {
            if (w.h == w.e) {
                if (!w.next()) break;
                qx = quad[w.ix];
                qy = quad[w.iy];
            }
            const PlasmaRec r = load(w.h++);
            const double charge_density_mu0 = r.q_invvol, vx = r.vx_c, vy = r.vy_c;
            const double gamma_psi = r.gamma_psi, psi_inv = r.psi_inv;
            const double tx = r.xmid, ty = r.ymid;
            const double tx2 = tx * tx, ty2 = ty * ty;
            const double sx_s = (qx.a2 * tx2 + qx.a1 * tx) + qx.a0;
            const double sx_ds = -((qx.b2 * tx2 + qx.b1 * tx) + qx.b0);
            const double sy_s = (qy.a2 * ty2 + qy.a1 * ty) + qy.a0;
            const double sy_ds = -((qy.b2 * ty2 + qy.b1 * ty) + qy.b0);
            if constexpr (SY) {
                acc += charge_density_mu0 * (
                    -sx_s * sy_s * (
                        -Bz_v * vx
                        + (Ez_v * vy
                        + ExmBy_v * (-vx * vy)
                        + EypBx_v * (gamma_psi - vy * vy)) * clight_inv
                        - 0.25 * AabssqDyp * q_mass_ratio * psi_inv
                    ) * q_mass_ratio * psi_inv
                    + (-sx_ds * sy_s * dxi * (-vx * vy)
                       - sx_s * sy_ds * dyi * (gamma_psi - vy * vy - 1.0)) * clight);
            } else {
                acc += charge_density_mu0 * (
                    +sx_s * sy_s * (
                        +Bz_v * vy
                        + (Ez_v * vx
                        + ExmBy_v * (gamma_psi - vx * vx)
                        + EypBx_v * (-vx * vy)) * clight_inv
                        - 0.25 * AabssqDxp * q_mass_ratio * psi_inv
                    ) * q_mass_ratio * psi_inv
                    + (+sx_ds * sy_s * dxi * (gamma_psi - vx * vx - 1.0)
                       + sx_s * sy_ds * dyi * (-vx * vy)) * clight);
            }
            ++hits;
        }
    };
    if (half == 0) run(std::integral_constant<int, 0>{});
    else run(std::integral_constant<int, 1>{});
    if (hits == 0) return;
    store[cout * ps + k] = acc;
}

// The fast pass of ::DepositCurrent for order 2: the per-cell sums of
// k_gather_current<2, true>, each cell walking its window as one sequence with the offset at
// run time (shape2_rt). The staged records hold t of x and y in place of the cell-scaled
// positions. The first thread of a cell sums jx, jy and jz, the second rho, chi and rhomjz.
__global__ __launch_bounds__(2 * GS * FLAT_BH) void k_gather_current_o2(
    Geom g, CLView cl, const double* __restrict__ rec, long n, double* __restrict__ store, int cjx, int cjy, int cjz,
    int crho, int cchi, int crm, double charge_mu0_mass_ratio, int32_t* __restrict__ ovf) {
    constexpr int O = 2, S = O + 1, DELTA = shape_delta<O>();
    __shared__ FlatStage st;
    auto& sl = st.l;
    auto copy = [&](int t, int pos) {
#pragma unroll
        for (int f = 0; f < 5; ++f) st.rec[f][flat_slot(t)] = rec[f * n + pos];
        const double xm = rec[5 * n + pos], ym = rec[6 * n + pos];
        st.rec[5][flat_slot(t)] = xm - floor(xm + 0.5);
        st.rec[6][flat_slot(t)] = ym - floor(ym + 0.5);
    };
    int b, i0, j0;
    if (!gather_setup<true, S, DELTA, FLAT_BH>(g, cl, sl, ovf, b, i0, j0, copy)) return;
    const int cell = int(threadIdx.x) % (GS * FLAT_BH), half = int(threadIdx.x) / (GS * FLAT_BH);
    const int li = cell % GS, lj = cell / GS;
    const int i = i0 + li, j = j0 + lj;
    const bool in = i <= g.nx - 1 + g.ng && j <= g.ny - 1 + g.ng;
    if (!block_raster_ordered_global<S, DELTA>(g, cl, in && half == 0, i, j, b, ovf)) return;
    if (!in) return;
    if (half == 0) crho = cchi = crm = -1;
    else cjx = cjy = cjz = -1;
    if (cjx < 0 && cjz < 0 && crho < 0 && cchi < 0 && crm < 0) return;
    const long k = g.idx(i, j);
    const long ps = g.plane;
    double ajx = cjx >= 0 ? store[cjx * ps + k] : 0.0, ajy = cjx >= 0 ? store[cjy * ps + k] : 0.0;
    double ajz = cjz >= 0 ? store[cjz * ps + k] : 0.0, arho = crho >= 0 ? store[crho * ps + k] : 0.0;
    double achi = cchi >= 0 ? store[cchi * ps + k] : 0.0, arm = crm >= 0 ? store[crm * ps + k] : 0.0;
    const double clight = 1.0;
    const double q_mu0_mass_ratio = charge_mu0_mass_ratio;
    const WindowColours<S> wc(g, cl, i, j);
    const FlatView src{sl, li, lj};
    FlatWalk<S, FlatView> w(wc, src, 0ull);
    double sgx = 0.0, sgy = 0.0;
    bool midx = false, midy = false;
    int hits = 0;
    while (true) {
        if (w.h == w.e) {
            if (!w.next()) break;
            sgx = w.ix == 0 ? -1.0 : 1.0;
            sgy = w.iy == 0 ? -1.0 : 1.0;
            midx = w.ix == 1;
            midy = w.iy == 1;
        }
        const int at = flat_slot(w.h++);
        const double q_invvol = st.rec[0][at], tx = st.rec[5][at], ty = st.rec[6][at];
        const double sx = shape2_rt(tx, sgx, midx);
        const double sy = shape2_rt(ty, sgy, midy);
        const double charge_density = q_invvol * sx * sy;
        if (cjx >= 0) {
            ajx += charge_density * st.rec[1][at];
            ajy += charge_density * st.rec[2][at];
        }
        if (cjz >= 0) ajz += charge_density * (st.rec[3][at] - 1.0) * clight;
        if (crho >= 0) arho += charge_density * st.rec[3][at];
        if (cchi >= 0) achi += charge_density * q_mu0_mass_ratio * st.rec[4][at];
        if (crm >= 0) arm += charge_density;
        ++hits;
    }
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

// ---- beam: per-cell contribution sequences -----------------------------------------------
// A beam's core puts hundreds of particles in a cell, in random index order, so merging a
// cell's window lists one element at a time would leave one thread per cell walking thousands
// of entries. Instead every (particle, stencil offset) pair finds its own place in its target
// cell's sequence at once: the sequence is the window's lists colour by colour, each colour
// merged by particle index, so a pair's place is the length of the lists of earlier colours
// plus, for every list of its own colour, the number of its entries with a smaller index (a
// binary search; the pair's own list gives its list rank). The pair writes its contribution
// there; then each cell adds its sequence in order to the plane's value, one thread per cell,
// or one wavefront per cell where the sequence is long.
constexpr int BEAM_LONG = 32;   // longer sequences: a wavefront per cell

// sequence length of every plane cell: the entries of its window lists (thread per cell)
template <int O>
__global__ __launch_bounds__(256) void k_beam_seg_count(Geom g, CLView cl, int32_t* __restrict__ seg) {
    constexpr int S = O + 1, DELTA = shape_delta<O>();
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t > g.plane) return;
    if (t == g.plane) {
        seg[t] = 0;
        return;
    }
    const int i = int(t % g.ldx) - g.ng, j = int(t / g.ldx) - g.ng;
    const int col0 = i - S + 1 - DELTA - cl.bxlo, row0 = j - S + 1 - DELTA - cl.bylo;
    int len = 0;
#pragma unroll
    for (int ky = 0; ky < S; ++ky) {
        const long r = long(row0 + ky) * cl.wx;
        len += cl.off[r + col0 + S] - cl.off[r + col0];
    }
    seg[t] = len;
}

// entries of perm[a, b) (increasing) below v
__device__ __forceinline__ int count_below(const int32_t* __restrict__ perm, int a, int b, int v) {
    int lo = a, hi = b;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (perm[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo - a;
}

// pair (list entry pos, stencil offset q): its place in the target cell's sequence and its
// contribution (BeamDepositCurrent.cpp:160-193: wq v times the two shape factors)
template <int O>
__global__ __launch_bounds__(256) void k_beam_pairs(Geom g, CLView cl, const int32_t* __restrict__ key,
                                                    const double* __restrict__ rec, long n,
                                                    const int32_t* __restrict__ seg, double* __restrict__ contrib,
                                                    long m) {
    constexpr int S = O + 1, DELTA = shape_delta<O>();
    const long tid = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long pl = tid / (S * S);
    const int q = int(tid - pl * (S * S));
    if (pl >= cl.off[long(cl.wx) * cl.wy]) return;
    const int pos = int(pl);
    const int p = cl.perm[pos];
    const int list = key[p];
    const int ix = q % S, iy = q / S;
    const int i = list % cl.wx + cl.bxlo + DELTA + ix, j = list / cl.wx + cl.bylo + DELTA + iy;
    if (i < -g.ng || i > g.nx - 1 + g.ng || j < -g.ng || j > g.ny - 1 + g.ng) return;
    const WindowColours<S> wc(g, cl, i, j);
    const int kp = (S - 1 - iy) * S + (S - 1 - ix);
    int cp = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c)
        if ((wc.mask[c] >> kp) & 1ull) cp = c;
    const int col0 = i - S + 1 - DELTA - cl.bxlo, row0 = j - S + 1 - DELTA - cl.bylo;
    int rank = pos - cl.off[list];
#pragma unroll
    for (int k = 0; k < S * S; ++k) {
        if (k == kp) continue;
        int c = 0;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc)
            if ((wc.mask[cc] >> k) & 1ull) c = cc;
        if (c > cp) continue;
        const long L = long(row0 + k / S) * cl.wx + col0 + k % S;
        const int a = cl.off[L], e = cl.off[L + 1];
        rank += c < cp ? e - a : count_below(cl.perm, a, e, p);
    }
    const long out = long(seg[g.idx(i, j)]) + rank;
    const double sx = rec[(3 + ix) * n + pos], sy = rec[(3 + O + 1 + iy) * n + pos];
    contrib[out] = sx * sy * rec[pos];
    contrib[m + out] = sx * sy * rec[n + pos];
    contrib[2 * m + out] = sx * sy * rec[2 * n + pos];
}

// the sums of cells with short sequences (thread per cell); longer ones are listed in longc
// (count at longc[0]) for k_beam_sum_long. jx and jy go together, as in the deposit functor.
__global__ __launch_bounds__(256) void k_beam_sum_short(Geom g, const int32_t* __restrict__ seg,
                                                        const double* __restrict__ contrib, long m,
                                                        double* __restrict__ store, int cjx, int cjy, int cjz,
                                                        int32_t* __restrict__ longc) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= g.plane) return;
    const int a = seg[t], e = seg[t + 1];
    if (a == e) return;
    if (e - a > BEAM_LONG) {
        longc[1 + atomicAdd(&longc[0], 1)] = int32_t(t);
        return;
    }
    const long ps = g.plane;
    if (cjx >= 0) {
        double ajx = store[cjx * ps + t], ajy = store[cjy * ps + t];
        for (int h = a; h < e; ++h) {
            ajx += contrib[h];
            ajy += contrib[m + h];
        }
        store[cjx * ps + t] = ajx;
        store[cjy * ps + t] = ajy;
    }
    if (cjz >= 0) {
        double ajz = store[cjz * ps + t];
        for (int h = a; h < e; ++h) ajz += contrib[2 * m + h];
        store[cjz * ps + t] = ajz;
    }
}

// the sums of the listed cells, a wavefront per cell: chunks of 64 contributions per component
// are read by the whole wavefront into LDS (the next chunk already in flight), and lanes 0-2
// add the jx, jy and jz chunks in sequence order
__global__ __launch_bounds__(64) void k_beam_sum_long(Geom g, const int32_t* __restrict__ seg,
                                                      const double* __restrict__ contrib, long m,
                                                      double* __restrict__ store, int cjx, int cjy, int cjz,
                                                      const int32_t* __restrict__ longc) {
    // LCH chunks of 64 contributions in flight: each chunk is read by the whole wavefront while
    // three lanes add the previous ones in sequence order
    constexpr int LCH = 4;
    __shared__ double buf[3][64];
    const int lane = int(threadIdx.x);
    const long ps = g.plane;
    const int comp = lane == 0 ? cjx : lane == 1 ? (cjx >= 0 ? cjy : -1) : lane == 2 ? cjz : -1;
    const int nl = longc[0];
    for (int w = int(blockIdx.x); w < nl; w += int(gridDim.x)) {
        const long t = longc[1 + w];
        const int a = seg[t], e = seg[t + 1];
        double acc = comp >= 0 ? store[comp * ps + t] : 0.0;
        double v[LCH][3];
#pragma unroll
        for (int d = 0; d < LCH; ++d) {
            const int h = a + 64 * d + lane;
            v[d][0] = v[d][1] = v[d][2] = 0.0;
            if (h < e) {
                v[d][0] = contrib[h];
                v[d][1] = contrib[m + h];
                v[d][2] = contrib[2 * m + h];
            }
        }
        for (int h0 = a; h0 < e; h0 += 64 * LCH) {
#pragma unroll
            for (int d = 0; d < LCH; ++d) {
                const int hc = h0 + 64 * d;
                if (hc >= e) break;
                __syncthreads();
                buf[0][lane] = v[d][0];
                buf[1][lane] = v[d][1];
                buf[2][lane] = v[d][2];
                __syncthreads();
                const int hn = hc + 64 * LCH + lane;
                if (hn < e) {
                    v[d][0] = contrib[hn];
                    v[d][1] = contrib[m + hn];
                    v[d][2] = contrib[2 * m + hn];
                }
                if (comp >= 0) {
                    // in sequence order; sixteen LDS reads in flight ahead of the adds
                    const int cnt = min(64, e - hc);
                    const double* row = buf[lane];
                    int u = 0;
                    for (; u + 16 <= cnt; u += 16) {
                        double q[16];
#pragma unroll
                        for (int k = 0; k < 16; ++k) q[k] = row[u + k];
#pragma unroll
                        for (int k = 0; k < 16; ++k) acc += q[k];
                    }
                    for (; u < cnt; ++u) acc += row[u];
                }
            }
        }
        if (comp >= 0) store[comp * ps + t] = acc;
        __syncthreads();
    }
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
long gather_blocks(const Geom& g, int bh) { return long((g.ldx + GS - 1) / GS) * ((g.ldy + bh - 1) / bh); }

// stride of the two overflow lists (the blocks of height FLAT_BH are the most numerous)
long ovf_stride(const Geom& g) { return gather_blocks(g, FLAT_BH) + 1; }

void lists_reserve(const Geom& g, CellLists& L, long n, int nrec) {
    const long ns = long(L.wx) * L.wy;
    // overflow lists of the gathers, [count, blocks...] each: the current (or beam)
    // deposition's at ovf, the explicit deposition's at ovf + gather_blocks + 1
    grow(L.ovf, L.ovf_cap, 2 * ovf_stride(g));
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

void lists_clear(const Geom& g, hipStream_t s, CellLists& L, int32_t* longc = nullptr) {
    const long ns = long(L.wx) * L.wy;
    hipLaunchKernelGGL(k_lists_clear, grid1d(ns + 1, 256), dim3(256), 0, s, L.off, ns + 1, L.ovf, ovf_stride(g),
                       longc);
    L.explicit_ovf_clear = true;
}

dim3 gather_grid(const Geom& g, int bh) { return dim3(unsigned(gather_blocks(g, bh))); }

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
                    (void*)L.rec, (void*)L.ovf, (void*)L.seg, (void*)L.segtmp, (void*)L.longc, (void*)L.contrib})
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
        // the fast pass, then the blocks it listed (merged, in the fast pass's block shape)
        if constexpr (o == 2) {
            hipLaunchKernelGGL(k_gather_current_o2, gather_grid(g, FLAT_BH), dim3(2 * GS * FLAT_BH), 0, c->stream, g,
                               view(L), L.rec, p->n, store, comps[0], comps[1], comps[2], comps[3], comps[4], comps[5],
                               cmr, L.ovf);
            hipLaunchKernelGGL((k_gather_current<o, false, FLAT_BH>), gather_grid(g, FLAT_BH), dim3(GS * FLAT_BH), 0,
                               c->stream, g, view(L), L.rec, p->n, store, comps[0], comps[1], comps[2], comps[3],
                               comps[4], comps[5], cmr, L.ovf);
        } else {
            hipLaunchKernelGGL((k_gather_current<o, true>), gather_grid(g, GS), dim3(plasma_gather_threads<true>()), 0,
                               c->stream, g, view(L), L.rec, p->n, store, comps[0], comps[1], comps[2], comps[3],
                               comps[4], comps[5], cmr, L.ovf);
            hipLaunchKernelGGL((k_gather_current<o, false>), gather_grid(g, GS), dim3(GS * GS), 0, c->stream, g,
                               view(L), L.rec, p->n, store, comps[0], comps[1], comps[2], comps[3], comps[4], comps[5],
                               cmr, L.ovf);
        }
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
            int32_t* ovf = L.ovf + ovf_stride(g);
            if (!L.explicit_ovf_clear) hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, ovf, 1L);
            L.explicit_ovf_clear = false;
            if constexpr (t == 2 && o == 2) {
                hipLaunchKernelGGL(k_gather_explicit_c2, gather_grid(g, FLAT_BH), dim3(2 * GS * FLAT_BH), 0, c->stream, g,
                                   view(L), L.rec, p->n, store, Ly.Bz, Ly.Ez, Ly.ExmBy, Ly.EypBx, Ly.Sy, Ly.Sx, qmr, ovf);
                hipLaunchKernelGGL((k_gather_explicit<t, o, false, FLAT_BH>), gather_grid(g, FLAT_BH),
                                   dim3(GS * FLAT_BH), 0, c->stream, g, view(L), L.rec, p->n, store, Ly.Bz, Ly.Ez,
                                   Ly.ExmBy, Ly.EypBx, Ly.Sy, Ly.Sx, qmr, ovf);
            } else {
                hipLaunchKernelGGL((k_gather_explicit<t, o, true>), gather_grid(g, GS), dim3(plasma_gather_threads<true>()),
                                   0, c->stream, g, view(L), L.rec, p->n, store, Ly.Bz, Ly.Ez, Ly.ExmBy, Ly.EypBx, Ly.Sy,
                                   Ly.Sx, qmr, ovf);
                hipLaunchKernelGGL((k_gather_explicit<t, o, false>), gather_grid(g, GS), dim3(GS * GS), 0, c->stream, g,
                                   view(L), L.rec, p->n, store, Ly.Bz, Ly.Ez, Ly.ExmBy, Ly.EypBx, Ly.Sy, Ly.Sx, qmr, ovf);
            }
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
    const int S = c->p.depos_order + 1;
    if (count > INT32_MAX / (S * S)) throw HpqError(HPQ_E_INVALID, "ordered beam deposition: too many particles in a slice");
    lists_geometry(c, L);
    lists_reserve(g, L, count, BEAM_NREC);
    grow(L.seg, L.seg_cap, g.plane + 1);
    grow(L.segtmp, L.segtmp_cap, scan_tmp_size(g.plane + 1));
    grow(L.longc, L.longc_cap, g.plane + 1);
    const long m = long(S) * S * count;
    grow(L.contrib, L.contrib_cap, 3 * m);
    TraceScope ts(c, HPQ_TRACE_BEAM_DEPOSIT);
    lists_clear(g, c->stream, L, L.longc);
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        hipLaunchKernelGGL(k_keys_beam<o>, grid1d(count, 256), dim3(256), 0, c->stream, g, *b, first, count, view(L),
                           select, select_value, L.key, L.slot, L.off);
        HPQ_HIP(hipGetLastError());
        lists_scatter(c->stream, L, count);
        hipLaunchKernelGGL(k_lists_order_beam<o>, grid1d(count, 256), dim3(256), 0, c->stream, g, *b, first, count,
                           L.key, L.off, L.unsorted, L.perm, L.rec);
        hipLaunchKernelGGL(k_beam_seg_count<o>, grid1d(g.plane + 1, 256), dim3(256), 0, c->stream, g, view(L), L.seg);
        scan_exclusive(c->stream, L.seg, g.plane + 1, L.segtmp);
        hipLaunchKernelGGL(k_beam_pairs<o>, grid1d(m, 256), dim3(256), 0, c->stream, g, view(L), L.key, L.rec, count,
                           L.seg, L.contrib, m);
    });
    hipLaunchKernelGGL(k_beam_sum_short, grid1d(g.plane, 256), dim3(256), 0, c->stream, g, L.seg, L.contrib, m, store,
                       comps[0], comps[1], comps[2], L.longc);
    hipLaunchKernelGGL(k_beam_sum_long, dim3(2048), dim3(64), 0, c->stream, g, L.seg, L.contrib, m, store, comps[0],
                       comps[1], comps[2], L.longc);
    HPQ_HIP(hipGetLastError());
}

}  // namespace hpq
