// Deterministic sorts of integer keys (cell, tile or slab indices).
//
// Replaces the reference's atomic linked-list cell build (CellLists/gpu.jl:28-61,
// whose intra-cell order depends on thread timing) and its target counting sort
// (CellLists/gpu.jl:183-219). Output order is (key, original index), so sums
// taken in sorted order are reproducible run to run.
//
//   1. histogram with atomics, remembering each element's arrival rank
//   2. exclusive scan of the histogram (multi-level, 1024 elements per block)
//   3. scatter to start[key] + rank
//   4. one thread per bin sorts its (short) segment by original index (skipped for the
//      pair kernel's target order, which no result depends on)
//
// radix_sort_bins gives the same order with a stable LSD radix sort of the keys
// (original indices as values) and derives the bin starts from the sorted keys;
// it suits bins too large for step 4.
#include "vpbs_internal.hpp"

#include <hipcub/hipcub.hpp>

namespace vpbs {

namespace {

constexpr int SCAN_BS = 256;
constexpr int SCAN_IPT = 4;
constexpr int SCAN_TILE = SCAN_BS * SCAN_IPT;

__global__ void fill_i32(int32_t *__restrict__ a, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void histogram_rank(const int32_t *__restrict__ keys, int64_t n, int32_t *__restrict__ counts,
                               int32_t *__restrict__ rank) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) rank[j] = atomicAdd(&counts[keys[j]], 1);
}

// In-place exclusive scan of one 1024-element tile per block; the tile total goes to sums[blockIdx].
__global__ __launch_bounds__(SCAN_BS) void scan_tiles(int32_t *__restrict__ a, int64_t n,
                                                      int32_t *__restrict__ sums) {
    __shared__ int32_t warp_tot[SCAN_BS / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IPT;
    int32_t v[SCAN_IPT];
    int32_t local = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        int64_t i = base + k;
        v[k] = i < n ? a[i] : 0;
        local += v[k];
    }
    // inclusive scan of `local` across the wavefront
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int32_t incl = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) warp_tot[wid] = incl;
    __syncthreads();
    int32_t wpre = 0;
    for (int w = 0; w < wid; w++) wpre += warp_tot[w];
    int32_t run = wpre + incl - local;  // exclusive prefix of this thread
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        int64_t i = base + k;
        if (i < n) a[i] = run;
        run += v[k];
    }
    if (threadIdx.x == SCAN_BS - 1) {
        int32_t tot = 0;
        for (int w = 0; w < SCAN_BS / 64; w++) tot += warp_tot[w];
        sums[blockIdx.x] = tot;
    }
}

__global__ void add_tile_offsets(int32_t *__restrict__ a, int64_t n, const int32_t *__restrict__ offs) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] += offs[i / SCAN_TILE];
}

__global__ void scatter_ranked(const int32_t *__restrict__ keys, const int32_t *__restrict__ rank, int64_t n,
                               const int32_t *__restrict__ start, int32_t *__restrict__ perm) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) perm[start[keys[j]] + rank[j]] = (int32_t)j;
}

__global__ void sort_segments(const int32_t *__restrict__ start, int64_t nbins, int32_t *__restrict__ perm) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    const int32_t s0 = start[b], s1 = start[b + 1];
    for (int32_t i = s0 + 1; i < s1; i++) {
        int32_t x = perm[i];
        int32_t k = i - 1;
        while (k >= s0 && perm[k] > x) {
            perm[k + 1] = perm[k];
            k--;
        }
        perm[k + 1] = x;
    }
}

__global__ void iota_i32(int32_t *__restrict__ a, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (int32_t)i;
}

// start[b] for every bin b ∈ [0, nbins] from the sorted keys: thread i ∈ [0, n] sets
// the bins in (key[i-1], key[i]] to i (key[-1] = -1, key[n] = nbins), so each bin
// is written exactly once.
__global__ void bin_starts(const int32_t *__restrict__ skeys, int64_t n, int64_t nbins, int32_t *__restrict__ start) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int64_t lo = i > 0 ? skeys[i - 1] : -1;
    const int64_t hi = i < n ? skeys[i] : nbins;
    for (int64_t b = lo + 1; b <= hi; b++) start[b] = (int32_t)i;
}

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Storage needed by the multi-level scan of n elements.
int64_t scan_tmp_size(int64_t n) {
    int64_t tot = 0;
    while (n > 1) {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        tot += n;
    }
    return tot + 1;
}

void exclusive_scan(int32_t *a, int64_t n, int32_t *tmp, hipStream_t s) {
    if (n <= 0) return;
    int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    scan_tiles<<<(unsigned)nt, SCAN_BS, 0, s>>>(a, n, tmp);
    if (nt > 1) {
        exclusive_scan(tmp, nt, tmp + nt, s);
        add_tile_offsets<<<nblocks(n, 256), 256, 0, s>>>(a, n, tmp);
    }
}

}  // namespace

int counting_sort(vpbs_ctx *c, SortBuffers &sb, const int32_t *keys, int64_t n, int64_t nbins, int32_t *start,
                  int32_t *perm, hipStream_t s, bool within_bin) {
    VPBS_TRY(grow(c, &sb.rank, &sb.cap_n, n));
    VPBS_TRY(grow(c, &sb.scan_tmp, &sb.cap_tmp, scan_tmp_size(nbins + 1)));
    fill_i32<<<nblocks(nbins + 1, 256), 256, 0, s>>>(start, nbins + 1, 0);
    if (n > 0) histogram_rank<<<nblocks(n, 256), 256, 0, s>>>(keys, n, start, sb.rank);
    exclusive_scan(start, nbins + 1, sb.scan_tmp, s);
    if (n > 0) scatter_ranked<<<nblocks(n, 256), 256, 0, s>>>(keys, sb.rank, n, start, perm);
    if (within_bin) sort_segments<<<nblocks(nbins, 256), 256, 0, s>>>(start, nbins, perm);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

int radix_sort_bins(vpbs_ctx *c, SortBuffers &sb, const int32_t *keys, int64_t n, int64_t nbins, int32_t *start,
                    int32_t *perm, hipStream_t s) {
    if (n > INT32_MAX || nbins >= INT32_MAX) return fail(c, VPBS_ERR_INVALID, "radix sort: too many elements or bins");
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < nbins) bits++;
    if (n > 0) {
        VPBS_TRY(grow(c, &sb.keys_out, &sb.cap_radix, 2 * n));  // sorted keys, then the original indices
        int32_t *iota = sb.keys_out + n;
        size_t tmp = 0;
        VPBS_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, sb.keys_out, iota, perm, (int)n, 0, bits, s));
        VPBS_TRY(grow(c, &sb.radix_tmp, &sb.cap_radix_tmp, (int64_t)tmp));
        iota_i32<<<nblocks(n, 256), 256, 0, s>>>(iota, n);
        tmp = (size_t)sb.cap_radix_tmp;
        VPBS_HIP(c, hipcub::DeviceRadixSort::SortPairs(sb.radix_tmp, tmp, keys, sb.keys_out, iota, perm, (int)n, 0, bits, s));
    }
    bin_starts<<<nblocks(n + 1, 256), 256, 0, s>>>(sb.keys_out, n, nbins, start);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

}  // namespace vpbs
