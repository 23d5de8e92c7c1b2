// Tile binning of plasma particles on gfx950: a stable counting sort by the tile key of
// the reference's CPU tiled deposition,
//   key = ((cx - lo) / T) * ntile_y + (cy - lo) / T,   invalid particles -> ntile_x * ntile_y
// (src/particles/deposition/DepositionUtil.H:213-227, amrex::DenseBins::build), where
// (cx, cy) is the lowest cell the particle deposits into (the get_start_cell functor,
// PlasmaDepositCurrent.cpp:137-155) and lo = -2 is the first guard cell of the slice box.
// The permutation lists particle indices bin by bin, in increasing index order inside a
// bin, which is numpy.argsort(key, kind="stable").
//
// Passes: keys (one thread per particle); per-chunk histograms in LDS, stored bin-major
// (bin b, chunk k at b * nchunks + k); exclusive scan of that matrix, which gives every
// (bin, chunk) pair its first output slot; scatter, one wavefront per chunk walking its
// particles in index order, so ties keep their order.
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"
#include "shape.hpp"

namespace hpq {

namespace {

constexpr int SCAN_T = 256, SCAN_ITEMS = 16, SCAN_BLOCK = SCAN_T * SCAN_ITEMS;
// bins of one counting sort: its per-chunk histograms live in LDS (128 KiB at this size)
constexpr long SORT_MAX_BINS = 32768;

// split > 1 (the atomic deposition's walk order): each tile's particles in `split` runs by
// particle index modulo split, so the particles a lattice injection put in one cell fall in
// different runs and neighbouring lanes of a deposition kernel add into different cells
template <int O>
__global__ __launch_bounds__(256) void k_bin_keys(Geom g, hpq_plasma p, int T, int TY, int ntx, int nty, int split,
                                                  int32_t* __restrict__ keys) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= p.n) return;
    int key = ntx * nty * split;
    if (p.valid[ip]) {
        const double xmid = (p.x[ip] - g.xoff) * g.dxi;
        const double ymid = (p.y[ip] - g.yoff) * g.dyi;
        const int cx = shape<O>(xmid, 0).cell, cy = shape<O>(ymid, 0).cell;
        // integer division truncates toward zero, as in the reference; the tile index is
        // clamped so a particle outside the guarded box still lands in a real bin
        const int tx = min(max((cx + g.ng) / T, 0), ntx - 1);
        const int ty = min(max((cy + g.ng) / TY, 0), nty - 1);
        key = (tx * nty + ty) * split + int(ip % split);
    }
    keys[ip] = key;
}

// per-chunk histogram, written bin-major
__global__ __launch_bounds__(256) void k_bin_count(const int32_t* __restrict__ keys, long n, int nbins, long chunk,
                                                   int nchunks, int32_t* __restrict__ counts) {
    extern __shared__ int hist[];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    const long lo = long(blockIdx.x) * chunk, hi = min(n, lo + chunk);
    for (long ip = lo + threadIdx.x; ip < hi; ip += blockDim.x) atomicAdd(&hist[keys[ip]], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) counts[long(b) * nchunks + blockIdx.x] = hist[b];
}

// one wavefront per chunk: stable ranks inside the chunk from a per-bin running count
__global__ __launch_bounds__(64) void k_bin_scatter(const int32_t* __restrict__ keys, long n, int nbins, long chunk,
                                                    int nchunks, const int32_t* __restrict__ first,
                                                    int32_t* __restrict__ perm) {
    extern __shared__ int run[];
    for (int b = threadIdx.x; b < nbins; b += 64) run[b] = 0;
    __syncthreads();
    const int lane = threadIdx.x;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const long lo = long(blockIdx.x) * chunk, hi = min(n, lo + chunk);
    for (long base = lo; base < hi; base += 64) {
        const long ip = base + lane;
        const bool act = ip < hi;
        const int key = act ? keys[ip] : -1;
        // peers: the active lanes holding the same key (wave-uniform loop over distinct keys)
        unsigned long long todo = __ballot(act);
        unsigned long long peers = 0;
        while (todo) {
            const int l = __ffsll((long long)todo) - 1;
            const int k = __shfl(key, l);
            const unsigned long long m = __ballot(act && key == k);
            if (act && key == k) peers = m;
            todo &= ~m;
        }
        const int leader = act ? __ffsll((long long)peers) - 1 : 0;
        int start = 0;
        if (act && lane == leader) {
            start = run[key];
            run[key] = start + __popcll(peers);
        }
        start = __shfl(start, leader);
        __syncthreads();
        if (act) perm[first[long(key) * nchunks + blockIdx.x] + start + __popcll(peers & lt)] = int32_t(ip);
    }
}

// block-local exclusive scan of SCAN_BLOCK elements; block total to sums[blockIdx.x]
__global__ __launch_bounds__(SCAN_T) void k_scan_block(int32_t* __restrict__ a, long n, int32_t* __restrict__ sums) {
    __shared__ int wsum[SCAN_T / 64];
    const long base = long(blockIdx.x) * SCAN_BLOCK + long(threadIdx.x) * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int tot = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        v[q] = (base + q < n) ? a[base + q] : 0;
        tot += v[q];
    }
    // exclusive scan of the per-thread totals across the block
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int inc = tot;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wv; ++w) woff += wsum[w];
    int run = woff + inc - tot;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        if (base + q < n) a[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == SCAN_T - 1 && sums) sums[blockIdx.x] = woff + inc;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_add(int32_t* __restrict__ a, long n, const int32_t* __restrict__ offs) {
    const long base = long(blockIdx.x) * SCAN_BLOCK + long(threadIdx.x) * SCAN_ITEMS;
    const int o = offs[blockIdx.x];
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q)
        if (base + q < n) a[base + q] += o;
}

// The two-launch scan for up to SCAN_FLAT_MAX blocks: block totals, then each block scans its
// elements from the sum of the totals before it (read by the whole block). A thread's
// SCAN_ITEMS consecutive elements are read as four 16-byte loads when the array is aligned.
constexpr long SCAN_FLAT_MAX = 4096;
__device__ __forceinline__ void scan_load(const int32_t* __restrict__ a, long base, long n, bool vec,
                                          int (&v)[SCAN_ITEMS]) {
    if (vec && base + SCAN_ITEMS <= n) {
        const int4* p = reinterpret_cast<const int4*>(a + base);
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS / 4; ++q) {
            const int4 x = p[q];
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) v[q] = (base + q < n) ? a[base + q] : 0;
    }
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const int32_t* __restrict__ a, long n, int vec,
                                                       int32_t* __restrict__ sums) {
    __shared__ int wsum[SCAN_T / 64];
    const long base = long(blockIdx.x) * SCAN_BLOCK + long(threadIdx.x) * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    scan_load(a, base, n, vec != 0, v);
    int tot = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) tot += v[q];
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < SCAN_T / 64; ++w) t += wsum[w];
        sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(SCAN_T) void k_scan_down(int32_t* __restrict__ a, long n, int vec,
                                                     const int32_t* __restrict__ sums) {
    __shared__ int wsum[SCAN_T / 64];
    __shared__ int wpre[SCAN_T / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // the sum of the totals of the blocks before this one
    int pre = 0;
    for (int k = threadIdx.x; k < int(blockIdx.x); k += SCAN_T) pre += sums[k];
    for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o);
    if (lane == 0) wpre[wv] = pre;
    const long base = long(blockIdx.x) * SCAN_BLOCK + long(threadIdx.x) * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    scan_load(a, base, n, vec != 0, v);
    int tot = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) tot += v[q];
    int inc = tot;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    int off = 0;
    for (int w = 0; w < SCAN_T / 64; ++w) off += wpre[w];
    for (int w = 0; w < wv; ++w) off += wsum[w];
    int run = off + inc - tot;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        if (base + q < n) a[base + q] = run;
        run += v[q];
    }
}

// bin starts: offsets[b] = first slot of bin b (its chunk-0 entry), offsets[nbins] = n;
// all zero when there are no particles (first == nullptr)
__global__ void k_bin_offsets(const int32_t* __restrict__ first, int nbins, int nchunks, long n,
                              int32_t* __restrict__ offsets) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbins) offsets[b] = first ? first[long(b) * nchunks] : 0;
    if (b == nbins) offsets[b] = int32_t(n);
}

template <class T>
void grow(T*& p, long& cap, long need) {
    if (need <= cap) return;
    if (p) HPQ_HIP(hipFree(p));
    p = nullptr;
    HPQ_HIP(hipMalloc(&p, sizeof(T) * size_t(need)));
    cap = need;
}

}  // namespace

// exclusive scan of a[0..n) in place; tmp holds the block sums of every level
void scan_exclusive(hipStream_t s, int32_t* a, long n, int32_t* tmp) {
    const long nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb <= 1) {
        hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(SCAN_T), 0, s, a, n, nullptr);
        return;
    }
    if (nb <= SCAN_FLAT_MAX) {
        const int vec = (reinterpret_cast<uintptr_t>(a) & 15) == 0 ? 1 : 0;
        hipLaunchKernelGGL(k_scan_reduce, dim3(unsigned(nb)), dim3(SCAN_T), 0, s, a, n, vec, tmp);
        hipLaunchKernelGGL(k_scan_down, dim3(unsigned(nb)), dim3(SCAN_T), 0, s, a, n, vec, tmp);
        return;
    }
    hipLaunchKernelGGL(k_scan_block, dim3(unsigned(nb)), dim3(SCAN_T), 0, s, a, n, tmp);
    scan_exclusive(s, tmp, nb, tmp + nb);
    hipLaunchKernelGGL(k_scan_add, dim3(unsigned(nb)), dim3(SCAN_T), 0, s, a, n, tmp);
}

long scan_tmp_size(long n) {
    long t = 0;
    for (long nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK; nb > 1; nb = (nb + SCAN_BLOCK - 1) / SCAN_BLOCK) t += nb;
    return t + 1;
}

void bins_free(hpq_ctx* c) {
    Bins& B = c->bins;
    for (void* q : {(void*)B.keys, (void*)B.perm, (void*)B.counts, (void*)B.tmp, (void*)B.reorder})
        if (q) (void)hipFree(q);
    B = Bins{};
}

void bin_tiles(hpq_ctx* c, const hpq_plasma* p, int tile, int32_t* perm, int32_t* offsets) {
    bin_tiles_rect(c, p, tile, tile, 1, perm, offsets);
}

// tiles of tile x tile_y cells, each split in `split` runs by particle index modulo split
void bin_tiles_rect(hpq_ctx* c, const hpq_plasma* p, int tile, int tile_y, int split, int32_t* perm,
                    int32_t* offsets) {
    if (tile <= 0 || tile_y <= 0 || split <= 0) throw HpqError(HPQ_E_INVALID, "bin_tiles: tile size must be positive");
    if (p->n > INT32_MAX) throw HpqError(HPQ_E_INVALID, "bin_tiles: more than 2^31-1 particles");
    const Geom& g = c->g;
    const int ntx = (g.nx + 2 * g.ng + tile - 1) / tile, nty = (g.ny + 2 * g.ng + tile_y - 1) / tile_y;
    if (long(ntx) * nty * split + 1 > SORT_MAX_BINS)
        throw HpqError(HPQ_E_UNSUPPORTED, split > 1 ? "bin_tiles: more than 32767 tile runs"
                                                    : "bin_tiles: more than 16383 tiles");
    if (split == 1 && ntx * nty + 1 > 16384) throw HpqError(HPQ_E_UNSUPPORTED, "bin_tiles: more than 16383 tiles");
    const int nbins = ntx * nty * split + 1;
    const long n = p->n;
    if (n == 0) {
        if (offsets)
            hipLaunchKernelGGL(k_bin_offsets, grid1d(nbins + 1, 256), dim3(256), 0, c->stream, nullptr, nbins, 1, 0L,
                               offsets);
        HPQ_HIP(hipGetLastError());
        return;
    }
    // at most 4096 chunks of at least 2048 particles (one wavefront walks a chunk)
    Bins& B = c->bins;
    grow(B.keys, B.keys_cap, n);
    dispatch_order(c->p.depos_order, [&](auto O) {
        hipLaunchKernelGGL((k_bin_keys<decltype(O)::value>), grid1d(n, 256), dim3(256), 0, c->stream, g, *p, tile,
                           tile_y, ntx, nty, split, B.keys);
    });
    HPQ_HIP(hipGetLastError());
    sort_keys(c, n, nbins, perm, offsets);
}

// chunking of the counting sort: at most 4096 chunks of at least 2048 particles (one
// wavefront walks a chunk)
static void sort_chunks(long n, long& chunk, int& nchunks) {
    chunk = std::max(2048L, (n + 4095) / 4096);
    nchunks = int((n + chunk - 1) / chunk);
}

// stable counting sort of keys[0..n) (values in [0, nbins)) with caller-provided work
// arrays: counts (nbins * nchunks entries), tmp (scan_tmp_size of that); writes the
// permutation (sorted slot -> index) and, when offsets != null, the first slot of every bin
// plus offsets[nbins] = n
static void sort_keys_into(hipStream_t s, const int32_t* keys, long n, int nbins, int32_t* counts, int32_t* tmp,
                           int32_t* perm, int32_t* offsets) {
    long chunk;
    int nchunks;
    sort_chunks(n, chunk, nchunks);
    const long m = long(nbins) * nchunks;
    const size_t lds = sizeof(int) * size_t(nbins);
    if (lds > 64 * 1024) {
        static std::once_flag once;
        std::call_once(once, [] {
            HPQ_HIP(hipFuncSetAttribute((const void*)k_bin_count, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        int(sizeof(int) * SORT_MAX_BINS)));
            HPQ_HIP(hipFuncSetAttribute((const void*)k_bin_scatter, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        int(sizeof(int) * SORT_MAX_BINS)));
        });
    }
    hipLaunchKernelGGL(k_bin_count, dim3(nchunks), dim3(256), lds, s, keys, n, nbins, chunk, nchunks, counts);
    scan_exclusive(s, counts, m, tmp);
    hipLaunchKernelGGL(k_bin_scatter, dim3(nchunks), dim3(64), lds, s, keys, n, nbins, chunk, nchunks, counts, perm);
    if (offsets)
        hipLaunchKernelGGL(k_bin_offsets, grid1d(nbins + 1, 256), dim3(256), 0, s, counts, nbins, nchunks, n,
                           offsets);
    HPQ_HIP(hipGetLastError());
}

// stable counting sort of the keys in bins.keys (n of them, values in [0, nbins)), work
// arrays held by the context
void sort_keys(hpq_ctx* c, long n, int nbins, int32_t* perm, int32_t* offsets) {
    Bins& B = c->bins;
    long chunk;
    int nchunks;
    sort_chunks(n, chunk, nchunks);
    const long m = long(nbins) * nchunks;
    grow(B.counts, B.counts_cap, m);
    grow(B.tmp, B.tmp_cap, scan_tmp_size(m));
    sort_keys_into(c->stream, B.keys, n, nbins, B.counts, B.tmp, perm, offsets);
}

// ---------------------------------------------------------------------------------
// Tile binning without allocation (hpq_bin_tiles_noalloc): the same kernels, their work
// arrays (keys, per-chunk histograms, scan partials) in one file-scope device array. Every
// device holds its own copy of the array from the library's load to the end of the process;
// its address on a device is looked up once, when the first context is created there
// (bins_noalloc_setup), and only that address is handed to kernels. All calls on one device
// share the array: calls that overlap in time on one device (two streams, two host
// threads) race. Nothing in it needs clearing: every entry a call reads was written earlier
// in the same call.
// ---------------------------------------------------------------------------------
namespace {

constexpr long NOALLOC_MAX_PARTICLES = 1L << 22;   // keys
constexpr long NOALLOC_MAX_COUNTS = 1L << 22;      // bins x chunks of the counting sort
constexpr long NOALLOC_TMP = 4096;                 // >= scan_tmp_size(NOALLOC_MAX_COUNTS) = 1025
constexpr long NOALLOC_INTS = NOALLOC_MAX_PARTICLES + NOALLOC_MAX_COUNTS + NOALLOC_TMP;
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) int32_t scratch[NOALLOC_INTS];

std::mutex g_noalloc_mu;
int32_t* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

}  // namespace

void bins_noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        throw HpqError(HPQ_E_UNSUPPORTED, "bin_tiles_noalloc: device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw HpqError(HPQ_E_HIP, std::string("bin_tiles_noalloc: hipGetSymbolAddress of the scratch array failed: ") +
                                      hipGetErrorString(e));
    g_noalloc_ptr[device] = static_cast<int32_t*>(ptr);
}

void bin_tiles_noalloc(hpq_ctx* c, const hpq_plasma* p, int tile, int32_t* perm, int32_t* offsets) {
    if (tile <= 0) throw HpqError(HPQ_E_INVALID, "bin_tiles: tile size must be positive");
    const Geom& g = c->g;
    const int ntx = (g.nx + 2 * g.ng + tile - 1) / tile, nty = (g.ny + 2 * g.ng + tile - 1) / tile;
    const int nbins = ntx * nty + 1;
    if (nbins > 16384) throw HpqError(HPQ_E_UNSUPPORTED, "bin_tiles: more than 16383 tiles");
    const long n = p->n;
    if (n > NOALLOC_MAX_PARTICLES)
        throw HpqError(HPQ_E_INVALID, "bin_tiles_noalloc: more than 4194304 particles (the scratch array's limit); "
                                      "use hpq_bin_tiles");
    long chunk;
    int nchunks;
    sort_chunks(std::max(n, 1L), chunk, nchunks);
    if (long(nbins) * nchunks > NOALLOC_MAX_COUNTS)
        throw HpqError(HPQ_E_INVALID, "bin_tiles_noalloc: tiles x particle chunks exceed 4194304 (the scratch "
                                      "array's limit); use larger tiles or hpq_bin_tiles");
    int32_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        base = (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) ? g_noalloc_ptr[c->device] : nullptr;
    }
    if (!base) throw HpqError(HPQ_E_HIP, "bin_tiles_noalloc: the scratch array of this device was not set up");
    if (n == 0) {
        if (offsets)
            hipLaunchKernelGGL(k_bin_offsets, grid1d(nbins + 1, 256), dim3(256), 0, c->stream, nullptr, nbins, 1, 0L,
                               offsets);
        HPQ_HIP(hipGetLastError());
        return;
    }
    int32_t* keys = base;
    int32_t* counts = base + NOALLOC_MAX_PARTICLES;
    int32_t* tmp = counts + NOALLOC_MAX_COUNTS;
    dispatch_order(c->p.depos_order, [&](auto O) {
        hipLaunchKernelGGL((k_bin_keys<decltype(O)::value>), grid1d(n, 256), dim3(256), 0, c->stream, g, *p, tile,
                           tile, ntx, nty, 1, keys);
    });
    HPQ_HIP(hipGetLastError());
    sort_keys_into(c->stream, keys, n, nbins, counts, tmp, perm, offsets);
}

namespace {

// BoxSorter key (src/particles/sorting/BoxSort.cpp:20-60): the zeta-slice of every valid
// beam particle; outside the domain or invalid -> nz (sorted last, made invalid)
__global__ void k_slice_keys(Geom g, hpq_beam_state b, int32_t* __restrict__ keys) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= b.n) return;
    int key = g.nz;
    if (b.valid[ip]) {
        const double s = floor((b.z[ip] - g.zlo) * g.dzi);
        if (s >= 0.0 && s < double(g.nz)) key = int(s);
    }
    keys[ip] = key;
}

template <class T>
__global__ void k_gather(const T* __restrict__ src, T* __restrict__ dst, const int32_t* __restrict__ perm, long n) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[perm[t]];
}

// the plasma's eleven FP64 arrays and valid flags gathered into staging in the order perm,
// then copied back (PlasmaIdx order of hpq_plasma)
struct PlasmaArrays {
    double* r[11];
    int32_t* valid;
};
__device__ __forceinline__ PlasmaArrays plasma_arrays(const hpq_plasma& p) {
    return {{p.x, p.y, p.w, p.ux, p.uy, p.psi, p.x_prev, p.y_prev, p.ux_half_step, p.uy_half_step, p.psi_half_step},
            p.valid};
}
__global__ __launch_bounds__(256) void k_reorder_gather(hpq_plasma p, const int32_t* __restrict__ perm,
                                                        double* __restrict__ st) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= p.n) return;
    const PlasmaArrays a = plasma_arrays(p);
    const long s = perm[t];
#pragma unroll
    for (int k = 0; k < 11; ++k) st[k * p.n + t] = a.r[k][s];
    reinterpret_cast<int32_t*>(st + 11 * p.n)[t] = a.valid[s];
}
__global__ __launch_bounds__(256) void k_reorder_back(hpq_plasma p, const double* __restrict__ st) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= p.n) return;
    const PlasmaArrays a = plasma_arrays(p);
#pragma unroll
    for (int k = 0; k < 11; ++k) a.r[k][t] = st[k * p.n + t];
    a.valid[t] = reinterpret_cast<const int32_t*>(st + 11 * p.n)[t];
}

// particles sorted past the last slice are outside the domain: invalid, zero weight
__global__ void k_invalidate_tail(hpq_beam_state b, const int32_t* __restrict__ offsets, int nz) {
    const long t = long(offsets[nz]) + long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= b.n) return;
    b.valid[t] = 0;
    b.w[t] = 0.0;
}

}  // namespace

void sort_beam_slices(hpq_ctx* c, hpq_beam_state* b, int64_t* off_out) {
    const Geom& g = c->g;
    const long n = b->n;
    Bins& B = c->bins;
    const int nbins = g.nz + 1;
    std::vector<int32_t> offs(nbins + 1, 0);
    if (n > 0) {
        if (n > INT32_MAX) throw HpqError(HPQ_E_INVALID, "sort_beam_slices: more than 2^31-1 particles");
        grow(B.keys, B.keys_cap, n);
        hipLaunchKernelGGL(k_slice_keys, grid1d(n, 256), dim3(256), 0, c->stream, g, *b, B.keys);
        HPQ_HIP(hipGetLastError());
        int32_t *perm = nullptr, *d_off = nullptr;
        void* scratch = nullptr;
        HPQ_HIP(hipMalloc((void**)&perm, sizeof(int32_t) * n));
        HPQ_HIP(hipMalloc((void**)&d_off, sizeof(int32_t) * (nbins + 1)));
        HPQ_HIP(hipMalloc(&scratch, sizeof(double) * n));
        sort_keys(c, n, nbins, perm, d_off);
        double* reals[7] = {b->x, b->y, b->z, b->ux, b->uy, b->uz, b->w};
        for (double* r : reals) {
            hipLaunchKernelGGL(k_gather<double>, grid1d(n, 256), dim3(256), 0, c->stream, r, (double*)scratch, perm, n);
            HPQ_HIP(hipMemcpyAsync(r, scratch, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        }
        for (int32_t* q : {b->valid, b->nsubcycles}) {
            hipLaunchKernelGGL(k_gather<int32_t>, grid1d(n, 256), dim3(256), 0, c->stream, q, (int32_t*)scratch, perm, n);
            HPQ_HIP(hipMemcpyAsync(q, scratch, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, c->stream));
        }
        hipLaunchKernelGGL(k_invalidate_tail, grid1d(n, 256), dim3(256), 0, c->stream, *b, d_off, g.nz);
        HPQ_HIP(hipGetLastError());
        HPQ_HIP(hipMemcpyAsync(offs.data(), d_off, sizeof(int32_t) * (nbins + 1), hipMemcpyDeviceToHost, c->stream));
        HPQ_HIP(hipStreamSynchronize(c->stream));
        (void)hipFree(perm);
        (void)hipFree(d_off);
        (void)hipFree(scratch);
    }
    for (int s = 0; s <= g.nz; ++s) off_out[s] = offs[s];
}

// the permutation the deposition kernels walk: 32x8-cell tiles, refreshed every
// bin_period slices by the slice driver; none when the plasma is reordered in place
const int32_t* deposit_order(hpq_ctx* c, const hpq_plasma* p) {
    Bins& B = c->bins;
    if (c->p.bin_period <= 0 || p->n <= 0 || plasma_reordered(c)) return nullptr;
    if (B.perm_n != p->n || B.stale) {
        grow(B.perm, B.perm_cap, p->n);
        bin_tiles_rect(c, p, DEPOSIT_TILE_X, DEPOSIT_TILE_Y, std::max(1, c->inj_ppc), B.perm, nullptr);
        B.perm_n = p->n;
        B.stale = false;
    }
    return B.perm;
}

// PlasmaParticleContainer::ReorderParticles (src/particles/plasma/PlasmaParticleContainer.cpp:
// 197-208; SortParticlesForDeposition on GPU builds): the plasma arrays in place in the order
// of the deposition tiles (the tile sort of deposit_order), so the atomic depositions read
// consecutive particles from consecutive addresses
void reorder_plasma(hpq_ctx* c, const hpq_plasma* p) {
    const long n = p->n;
    if (n <= 0) return;
    Bins& B = c->bins;
    grow(B.perm, B.perm_cap, n);
    // from injection order the first reorder of a step splits each tile by a particle's index
    // within its cell (one particle per cell in every run of a tile, so the lanes of a
    // wavefront add to different LDS words); later ones keep the order inside a tile (stable)
    bin_tiles_rect(c, p, DEPOSIT_TILE_X, DEPOSIT_TILE_Y, B.reordered ? 1 : std::max(1, c->inj_ppc), B.perm, nullptr);
    B.perm_n = -1;   // the permutation describes the old order
    B.stale = true;
    B.reordered = true;
    // 11 FP64 arrays and the int32 flags (half a double each, rounded up)
    grow(B.reorder, B.reorder_cap, 11 * n + (n + 1) / 2);
    hipLaunchKernelGGL(k_reorder_gather, grid1d(n, 256), dim3(256), 0, c->stream, *p, B.perm, B.reorder);
    hipLaunchKernelGGL(k_reorder_back, grid1d(n, 256), dim3(256), 0, c->stream, *p, B.reorder);
    HPQ_HIP(hipGetLastError());
}

}  // namespace hpq
