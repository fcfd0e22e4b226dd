// Long-range part of the Ewald-split Biot–Savart sum on gfx950: a NUFFT with an
// "exponential of semicircle" (ES) kernel, rocFFT, and one fused Fourier-space
// kernel.
//
//   ω̂(k) = (Γ/V) Σ_j q_j e^{-ik·s_j}         type 1 (longrange.jl:488-495,
//                                           nonuniformffts.jl:173-192)
//   v̂ = i k × ω̂ / k²,  ψ̂ = ω̂ / k²          (longrange.jl:102-142)
//   × exp(-k²/4α²)                          Ewald Gaussian (cache_common.jl:64-76)
//   u(x_i) = Σ_k û(k) e^{ik·x_i}             type 2 (nonuniformffts.jl:195-212)
//
// restricted to the modes kept by _ensure_hermitian_symmetry!
// (longrange.jl:556-583), optionally to |k| ≤ kmax (truncate_spherical).
//
// Kernels:
//   lr_tile_keys    tile of each quadrature point on the σ-oversampled grid
//   spread_rows_kernel  one wavefront per grid tile (one 128-B row wide, 8×8 rows)
//                   takes the points of the 3×3×3 neighbouring tiles whose support
//                   overlaps it and adds their kernel-weighted charges into the tile
//                   rows held in registers (lane = row): every grid value written
//                   exactly once, no atomics, no zero-fill pass, deterministic order
//   FFT             batched contiguous 1D c2c transforms on rocFFT along x, y and
//                   z, with pencil transposes of our own between them:
//                   x: the real grid holds pairs of y-rows as one complex row, so
//                      one c2c transforms two real rows;
//                   pair_split_to_ypencils: separates the pair, A = (Z_k + conj Z_-k)/2,
//                      B = (Z_k - conj Z_-k)/2i, keeps the retained kx and makes
//                      y contiguous; ypencils_to_zpencils keeps the retained ky and
//                      makes z contiguous. The inverse mirrors this and rebuilds
//                      the Hermitian half kx < 0 and the zero modes.
//   fourier_kernel  deconvolution (twice: type 1 and type 2), Γ/V, Green function,
//                   curl, Ewald Gaussian, zeroing of dropped modes — one pass over
//                   the z-pencils of retained (kx, ky), in place
//   interp_kernel_wave  W³ ES-weighted gather, one wavefront per node
// Everything is templated on the grid type (double for the FP64 path, float
// for the FP32-bulk path).
#include "vpbs_internal.hpp"
#include "device_util.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

// Scratch of the allocation-free energy spectrum (vpbs_energy_spectrum_noalloc): the
// per-block partial spectra followed by the summed spectrum. Each GPU the library runs on
// holds its own copy until the process ends; host code reaches it only through the
// per-device address taken with hipGetSymbolAddress at context creation.
constexpr int SPEC_NA_BLOCKS = 256;
constexpr int SPEC_NA_MAXBINS = 1024;
static __device__ __align__(256) double scratch[(size_t)SPEC_NA_BLOCKS * SPEC_NA_MAXBINS + SPEC_NA_MAXBINS];

namespace vpbs {

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};  // device address of `scratch` on each GPU (null: not taken yet)


// Spreading tile: 16 × 8 × 8 cells. A pair row of a tile (rows y and y+1
// interleaved) is 32 values, 128 B in FP32; 16 x-cells keep the per-lane
// accumulators (16 × 3) in registers at 4 wavefronts per SIMD.
template <typename CT>
constexpr int tile_x() {
    return 16;
}
constexpr int TY = 8, TZ = 8;
constexpr int MAXM = 8;     // largest supported half-support

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T>
__device__ inline T es_kernel(T z, T beta) {
    T s = (T)1 - z * z;
    return s < (T)0 ? (T)0 : exp(beta * (sqrt(s) - (T)1));
}

__device__ inline int wrapi(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

// i in (-n, 2n) -> i mod n without an integer division
__device__ inline int wrap1(int i, int n) {
    i += (i < 0) ? n : 0;
    i -= (i >= n) ? n : 0;
    return i;
}

struct LRGeom {
    int G[3];
    double scale[3];  // G/L
    int ntile[3];
    int tx;           // spreading tile width along x (cells)
    int M;
    int64_t cstride;  // real grid values per component
};

// Per tile-sorted point, the record read by spreading: the 2m kernel weights
// along x, y and z, then the 3 charge components, padded to 16 bytes.
constexpr int rec_len(int W) { return (3 * W + 3 + 3) & ~3; }

// Position of real grid value (x, y, z) of one component: rows y = 2p and
// y = 2p + 1 are interleaved as the real and imaginary parts of complex row p.
__device__ inline int64_t gidx(const LRGeom &g, int x, int y, int z) {
    return (((int64_t)z * (g.G[1] >> 1) + (y >> 1)) * g.G[0] + x) * 2 + (y & 1);
}

template <typename CT>
struct alignas(2 * sizeof(CT)) Cx {
    CT re, im;
};

// ----------------------------------------------------------------------------

__global__ void lr_tile_keys(int64_t Nq, const double *__restrict__ pts, LRGeom g, int32_t *__restrict__ key) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= Nq) return;
    int t[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double u = pts[d * Nq + j] * g.scale[d];
        int i = wrapi((int)floor(u), g.G[d]);
        t[d] = i / (d == 0 ? g.tx : (d == 1 ? TY : TZ));
    }
    key[j] = (t[2] * g.ntile[1] + t[1]) * g.ntile[0] + t[0];
}

// Tile-sorted copy of the points: per point and direction d the first stencil
// index i0_d = ⌊u_d⌋ - m + 1 (wrapped into [0, G_d), component-major for the
// overlap tests) and one record (rec_len(2m) values) holding the 2m kernel
// weights φ((i0_d + k - u_d)/m) per direction and the charge, computed once
// here instead of once per tile in the spreading kernel.
// Records are rec_len(W) values (a multiple of 16 bytes) moved with 16-byte vector
// accesses: one thread writes or reads a whole record.
template <typename CT, int N>
__device__ __forceinline__ void store_rec(CT *__restrict__ p, const CT (&r)[N]) {
    static_assert((N * sizeof(CT)) % 16 == 0, "records are whole 16-byte vectors");
    if constexpr (sizeof(CT) == 4) {
#pragma unroll
        for (int k = 0; k < N; k += 4) reinterpret_cast<float4 *>(p)[k / 4] = float4{r[k], r[k + 1], r[k + 2], r[k + 3]};
    } else {
#pragma unroll
        for (int k = 0; k < N; k += 2) reinterpret_cast<double2 *>(p)[k / 2] = double2{r[k], r[k + 1]};
    }
}
template <typename CT, int N>
__device__ __forceinline__ void load_rec(const CT *__restrict__ p, CT (&r)[N]) {
    if constexpr (sizeof(CT) == 4) {
#pragma unroll
        for (int k = 0; k < N; k += 4) {
            const float4 v = reinterpret_cast<const float4 *>(p)[k / 4];
            r[k] = v.x, r[k + 1] = v.y, r[k + 2] = v.z, r[k + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; k += 2) {
            const double2 v = reinterpret_cast<const double2 *>(p)[k / 2];
            r[k] = v.x, r[k + 1] = v.y;
        }
    }
}

template <typename CT, int W>
__global__ void gather_points(int64_t Nq, const int32_t *__restrict__ perm, const double *__restrict__ pts,
                              const double *__restrict__ chg, LRGeom g, CT beta, int32_t *__restrict__ si0,
                              CT *__restrict__ rec) {
    constexpr int M = W / 2, REC = rec_len(W);
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Nq) return;
    const int32_t j = perm[s];
    CT r[REC];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double u = pts[d * Nq + j] * g.scale[d];
        const int i0 = (int)floor(u) - M + 1;
        si0[d * Nq + s] = wrapi(i0, g.G[d]);
#pragma unroll
        for (int k = 0; k < W; k++) r[d * W + k] = es_kernel<CT>((CT)(((double)(i0 + k) - u) / (double)M), beta);
        r[3 * W + d] = (CT)chg[d * Nq + j];
    }
#pragma unroll
    for (int k = 3 * W + 3; k < REC; k++) r[k] = (CT)0;
    store_rec(rec + s * REC, r);
}

template <typename CT>
void launch_gather(int64_t Nq, const int32_t *perm, const double *pts, const double *chg, const LRGeom &g, CT beta,
                   int32_t *si0, CT *rec, hipStream_t s) {
    const unsigned nb = (unsigned)((Nq + 255) / 256);
    switch (g.M) {
        case 2: gather_points<CT, 4><<<nb, 256, 0, s>>>(Nq, perm, pts, chg, g, beta, si0, rec); break;
        case 3: gather_points<CT, 6><<<nb, 256, 0, s>>>(Nq, perm, pts, chg, g, beta, si0, rec); break;
        case 4: gather_points<CT, 8><<<nb, 256, 0, s>>>(Nq, perm, pts, chg, g, beta, si0, rec); break;
        case 5: gather_points<CT, 10><<<nb, 256, 0, s>>>(Nq, perm, pts, chg, g, beta, si0, rec); break;
        case 6: gather_points<CT, 12><<<nb, 256, 0, s>>>(Nq, perm, pts, chg, g, beta, si0, rec); break;
        case 7: gather_points<CT, 14><<<nb, 256, 0, s>>>(Nq, perm, pts, chg, g, beta, si0, rec); break;
        default: gather_points<CT, 16><<<nb, 256, 0, s>>>(Nq, perm, pts, chg, g, beta, si0, rec); break;
    }
}

// Home slab of each node for the interpolation fused with the inverse x transform:
// the slab (SLAB_HY y-rows × SLAB_HZ z-rows) of its first stencil cell (y, z), the
// cell gather_nodes stores.
__global__ void slab_keys(int64_t n, const double *__restrict__ pos, LRGeom g, int32_t *__restrict__ key) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int i0[3];
#pragma unroll
    for (int d = 1; d < 3; d++) {
        const double u = pos[d * n + j] * g.scale[d];
        i0[d] = wrapi((int)floor(u) - g.M + 1, g.G[d]);
    }
    key[j] = (i0[2] / SLAB_HZ) * (g.G[1] / SLAB_HY) + i0[1] / SLAB_HY;
}

// Slab-sorted copy of the nodes: first stencil cell per direction (wrapped into
// [0, G_d), component-major) and one record of the W kernel weights per direction,
// φ((i0_d + k - u_d)/m) as interp_kernel_wave evaluates them, padded with zeros.
template <typename CT, int W>
__global__ void gather_nodes(int64_t Np, const int32_t *__restrict__ perm, const double *__restrict__ nodes, LRGeom g,
                             CT beta, int32_t *__restrict__ si0, CT *__restrict__ rec) {
    constexpr int M = W / 2, REC = slab_rec_len(W, (int)sizeof(CT));
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Np) return;
    const int32_t j = perm[s];
    CT r[REC];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double u = nodes[d * Np + j] * g.scale[d];
        const int i0 = (int)floor(u) - M + 1;
        si0[d * Np + s] = wrapi(i0, g.G[d]);
#pragma unroll
        for (int k = 0; k < W; k++) r[d * W + k] = es_kernel<CT>((CT)(((double)(i0 + k) - u) / (double)M), beta);
    }
#pragma unroll
    for (int k = 3 * W; k < REC; k++) r[k] = (CT)0;
    store_rec(rec + s * REC, r);
}

template <typename CT>
void launch_gather_nodes(int64_t Np, const int32_t *perm, const double *nodes, const LRGeom &g, CT beta, int32_t *si0,
                         CT *rec, hipStream_t s) {
    const unsigned nb = (unsigned)((Np + 255) / 256);
    switch (g.M) {
        case 2: gather_nodes<CT, 4><<<nb, 256, 0, s>>>(Np, perm, nodes, g, beta, si0, rec); break;
        case 3: gather_nodes<CT, 6><<<nb, 256, 0, s>>>(Np, perm, nodes, g, beta, si0, rec); break;
        case 4: gather_nodes<CT, 8><<<nb, 256, 0, s>>>(Np, perm, nodes, g, beta, si0, rec); break;
        case 5: gather_nodes<CT, 10><<<nb, 256, 0, s>>>(Np, perm, nodes, g, beta, si0, rec); break;
        case 6: gather_nodes<CT, 12><<<nb, 256, 0, s>>>(Np, perm, nodes, g, beta, si0, rec); break;
        case 7: gather_nodes<CT, 14><<<nb, 256, 0, s>>>(Np, perm, nodes, g, beta, si0, rec); break;
        default: gather_nodes<CT, 16><<<nb, 256, 0, s>>>(Np, perm, nodes, g, beta, si0, rec); break;
    }
}

// ----------------------------------------------------------------------------
// Spreading (type 1), gather form.

// active[t] = 1 if any of the 3×3×3 tiles around t holds a point (only those
// tiles can receive a non-zero value).
__global__ void tile_active(LRGeom g, int64_t ntiles, const int32_t *__restrict__ tstart, uint8_t *__restrict__ active) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int tx = (int)(t % g.ntile[0]), ty = (int)((t / g.ntile[0]) % g.ntile[1]);
    const int tz = (int)(t / ((int64_t)g.ntile[0] * g.ntile[1]));
    int any = 0;
    for (int c = -1; c <= 1 && !any; c++)
        for (int b = -1; b <= 1 && !any; b++)
            for (int a = -1; a <= 1; a++) {
                int64_t tl = ((int64_t)wrapi(tz + c, g.ntile[2]) * g.ntile[1] + wrapi(ty + b, g.ntile[1])) * g.ntile[0] +
                             wrapi(tx + a, g.ntile[0]);
                if (tstart[tl + 1] > tstart[tl]) {
                    any = 1;
                    break;
                }
            }
    active[t] = (uint8_t)any;
}

// Block b of the launch handles tile xcd_tile(b): blocks are dealt to the 8 XCDs
// round-robin, so consecutive blocks of one XCD get consecutive tiles and the
// point data shared by neighbouring tiles stays in that XCD's L2.
__device__ inline int64_t xcd_tile(int64_t b, int64_t ntiles) {
    const int64_t per = (ntiles + 7) / 8;
    return (b % 8) * per + b / 8;
}

template <typename CT>
struct alignas(16) Q4 {
    CT x, y, z, w;
};

// f(std::integral_constant<int, I>) for I = B .. E-1: register arrays indexed
// through it keep compile-time indices and stay in registers.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Spreading (type 1): one wavefront per TX×8×8 tile, lane (y, z) = (lane % 8,
// lane / 8) owning tile row (y, z) and accumulating its TX cells × 3
// components in registers.
//  1) lanes 0..26 read the point ranges of the 3×3×3 neighbour tiles at once;
//  2) the candidates are tested 64 at a time for support overlap; the kept ones
//     are compacted in order and their records (weights, charge) copied to LDS;
//  3) the wavefront walks the kept points: each lane weights the charge by its
//     row's wy·wz and adds w16[x]·(…) to all TX cells of the row, where w16 holds
//     the point's x-weights already shifted to the tile columns (zeros elsewhere,
//     written when the record is staged): branch-free, compile-time register
//     indices, every weight read is an LDS broadcast.
// Summation order is fixed (neighbour tile, then sorted order): results are
// deterministic. The tile is written once, transposed through LDS into
// contiguous pair rows.
template <typename CT, int W>
__global__ __launch_bounds__(64) void spread_rows_kernel(LRGeom g, int64_t Nq, const int32_t *__restrict__ tstart,
                                                         const uint8_t *__restrict__ active,
                                                         const int32_t *__restrict__ si0,
                                                         const CT *__restrict__ rec, CT *__restrict__ grid) {
    constexpr int TX = tile_x<CT>();
    constexpr int REC = rec_len(W);
    constexpr int TXP = TX + 1;  // LDS row pitch (odd: conflict-free column writes)
    constexpr int NPAIR = TY / 2;
    // staged kept points: y/z weights and charge (YZQ values), x-weights at their tile
    // columns (TX values); the write-out buffer reuses the same LDS afterwards
    constexpr int YZQ = 2 * W + 3;
    static_assert(TY * TZ * TXP <= 64 * (YZQ + TX), "write-out buffer fits the staging area");
    __shared__ CT s_stage[64 * (YZQ + TX)];
    CT(*s_rec)[YZQ] = reinterpret_cast<CT(*)[YZQ]>(s_stage);
    CT(*s_w16)[TX] = reinterpret_cast<CT(*)[TX]>(s_stage + 64 * YZQ);
    CT *buf = s_stage;
    __shared__ int4 s_meta[64];  // (sx, ry, rz, -) of the kept points
    __shared__ int nb_start[27];
    __shared__ int nb_pre[28];

    const int lane = threadIdx.x;
    const int64_t ntl = (int64_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
    const int64_t tile = xcd_tile(blockIdx.x, ntl);
    if (tile >= ntl) return;  // the whole (one-wavefront) block
    const int tix = (int)(tile % g.ntile[0]);
    const int tiy = (int)((tile / g.ntile[0]) % g.ntile[1]);
    const int tiz = (int)(tile / ((int64_t)g.ntile[0] * g.ntile[1]));
    const int t0x = tix * TX, t0y = tiy * TY, t0z = tiz * TZ;
    const int ly = lane & 7, lz = lane >> 3;

    CT acc[TX][3];
    static_for<0, TX>([&](auto X) {
        constexpr int x = decltype(X)::value;
        acc[x][0] = acc[x][1] = acc[x][2] = (CT)0;
    });

    if (active[tile]) {
        // 1) neighbour ranges; with fewer than 3 tiles along d a neighbour repeats: empty range
        int len = 0;
        if (lane < 27) {
            const int o[3] = {lane % 3 - 1, (lane / 3) % 3 - 1, lane / 9 - 1};
            const int tt[3] = {tix, tiy, tiz};
            int tn[3];
            bool dup = false;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                tn[d] = wrap1(tt[d] + o[d], g.ntile[d]);
                for (int op = -1; op < o[d]; op++) dup |= (wrap1(tt[d] + op, g.ntile[d]) == tn[d]);
            }
            const int64_t tl = ((int64_t)tn[2] * g.ntile[1] + tn[1]) * g.ntile[0] + tn[0];
            const int s0 = tstart[tl];
            nb_start[lane] = s0;
            len = dup ? 0 : tstart[tl + 1] - s0;
        }
        const int incl = wave_incl_scan(len);  // DPP steps, no LDS round trips
        if (lane < 27) nb_pre[lane + 1] = incl;
        if (lane == 0) nb_pre[0] = 0;
        const int ncand = __builtin_amdgcn_readlane(incl, 26);
        __syncthreads();

        for (int base = 0; base < ncand; base += 64) {
            // 2) overlap test, compaction, record staging
            const int cnd = base + lane;
            bool keep = false;
            int sidx = 0, sx = 0, ry = 0, rz = 0;
            if (cnd < ncand) {
                int lo = 0, hi = 26;  // largest neighbour nb with nb_pre[nb] <= cnd
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (nb_pre[mid] <= cnd)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                sidx = nb_start[lo] + (cnd - nb_pre[lo]);
                const int ix = si0[sidx], iy = si0[Nq + sidx], iz = si0[2 * Nq + sidx];
                const int rx = wrap1(ix - t0x, g.G[0]);  // support start relative to the tile start
                ry = wrap1(iy - t0y, g.G[1]);
                rz = wrap1(iz - t0z, g.G[2]);
                keep = (rx < TX || wrap1(t0x - ix, g.G[0]) < W) && (ry < TY || wrap1(t0y - iy, g.G[1]) < W) &&
                       (rz < TZ || wrap1(t0z - iz, g.G[2]) < W);
                sx = rx < TX ? rx : rx - g.G[0];  // ∈ (-W, TX)
                ry = ry < TY ? ry : ry - g.G[1];  // ∈ (-W, TY), likewise below
                rz = rz < TZ ? rz : rz - g.G[2];
            }
            // the window sx + [0, W) lies in the left half of the row when sx ≤ 0, in the
            // right half when sx ≥ TX/2 (W ≤ TX/2, more than one tile along x): the kept
            // points are compacted in three groups (left, right, both), each then walked
            // by its own branch-free loop
            constexpr int H = TX / 2;
            const int cls = (W <= H && sx <= 0 && g.G[0] != TX) ? 0 : (W <= H && sx >= H) ? 1 : 2;
            const unsigned long long m0 = __ballot(keep && cls == 0), m1 = __ballot(keep && cls == 1),
                                     m2 = __ballot(keep && cls == 2);
            const int n0 = __popcll(m0), n1 = __popcll(m1), nkept = n0 + n1 + __popcll(m2);
            if (keep) {
                const unsigned long long lt = (1ull << lane) - 1ull;
                const int pos = cls == 0 ? __popcll(m0 & lt) : cls == 1 ? n0 + __popcll(m1 & lt) : n0 + n1 + __popcll(m2 & lt);
                s_meta[pos] = make_int4(sx, ry, rz, 0);
                CT pr[REC];
                load_rec(rec + (int64_t)sidx * REC, pr);
#pragma unroll
                for (int k = 0; k < YZQ; k++) s_rec[pos][k] = pr[W + k];
#pragma unroll
                for (int x = 0; x < TX; x++) s_w16[pos][x] = (CT)0;
                // x-weights at their tile columns sx + k (mod TX when one tile spans the period)
#pragma unroll
                for (int k = 0; k < W; k++) {
                    int x = sx + k;
                    if (g.G[0] == TX) x = wrap1(x, TX);
                    if (x >= 0 && x < TX) s_w16[pos][x] = pr[k];
                }
            }
            __syncthreads();
            // 3) accumulate the kept points in order: acc[x] += w16[x]·(wy·wz·q) over the
            //    columns [X0, X1) of the row, w16 = the point's x-weights shifted to the
            //    tile (zero outside)
            auto walk = [&](int p0, int p1, auto X0, auto X1) {
                for (int p = p0; p < p1; p++) {
                    const int4 m = s_meta[p];
                    const int ky = ly - m.y, kz = lz - m.z;
                    CT wyz = (CT)0;
                    if ((unsigned)ky < (unsigned)W && (unsigned)kz < (unsigned)W)
                        wyz = s_rec[p][ky] * s_rec[p][W + kz];
                    const CT a0 = wyz * s_rec[p][2 * W], a1 = wyz * s_rec[p][2 * W + 1],
                             a2 = wyz * s_rec[p][2 * W + 2];
                    static_for<decltype(X0)::value, decltype(X1)::value>([&](auto X) {
                        constexpr int x = decltype(X)::value;
                        const CT w = s_w16[p][x];
                        acc[x][0] += w * a0;
                        acc[x][1] += w * a1;
                        acc[x][2] += w * a2;
                    });
                }
            };
            walk(0, n0, std::integral_constant<int, 0>{}, std::integral_constant<int, H>{});
            walk(n0, n0 + n1, std::integral_constant<int, H>{}, std::integral_constant<int, TX>{});
            walk(n0 + n1, nkept, std::integral_constant<int, 0>{}, std::integral_constant<int, TX>{});
            __syncthreads();
        }
    }
    // write: per component, rows to LDS, then contiguous pair rows (2·TX values) to the grid;
    // 32-bit offsets from the tile origin (a tile spans 8 z-planes)
    constexpr int VPR = 2 * TX;     // values per pair row
    constexpr int RPI = 64 / VPR;   // pair rows per store instruction
    CT *const tile0 = grid + (((int64_t)t0z * (g.G[1] >> 1) + (t0y >> 1)) * g.G[0] + t0x) * 2;
    const int rstride = 2 * g.G[0], zstride = (g.G[1] >> 1) * rstride;
    static_for<0, 3>([&](auto C) {
        constexpr int c = decltype(C)::value;
        static_for<0, TX>([&](auto X) {
            constexpr int x = decltype(X)::value;
            buf[(lz * TY + ly) * TXP + x] = acc[x][c];
        });
        __syncthreads();
        for (int it = 0; it < TZ * NPAIR / RPI; it++) {
            const int row = it * RPI + lane / VPR;  // z·NPAIR + pair
            const int v = lane % VPR;
            const int z = row / NPAIR, pp = row - z * NPAIR;
            const CT val = buf[(z * TY + 2 * pp + (v & 1)) * TXP + (v >> 1)];
            tile0[c * g.cstride + (z * zstride + pp * rstride + v)] = val;
        }
        __syncthreads();
    });
}

template <typename CT>
void launch_spread(const LRGeom &g, int64_t nb, int64_t Nq, const int32_t *tstart, const uint8_t *active,
                   const int32_t *si0, const CT *rec, CT *grid, hipStream_t s) {
    const unsigned n = (unsigned)nb;
    switch (g.M) {
        case 2: spread_rows_kernel<CT, 4><<<n, 64, 0, s>>>(g, Nq, tstart, active, si0, rec, grid); break;
        case 3: spread_rows_kernel<CT, 6><<<n, 64, 0, s>>>(g, Nq, tstart, active, si0, rec, grid); break;
        case 4: spread_rows_kernel<CT, 8><<<n, 64, 0, s>>>(g, Nq, tstart, active, si0, rec, grid); break;
        case 5: spread_rows_kernel<CT, 10><<<n, 64, 0, s>>>(g, Nq, tstart, active, si0, rec, grid); break;
        case 6: spread_rows_kernel<CT, 12><<<n, 64, 0, s>>>(g, Nq, tstart, active, si0, rec, grid); break;
        case 7: spread_rows_kernel<CT, 14><<<n, 64, 0, s>>>(g, Nq, tstart, active, si0, rec, grid); break;
        default: spread_rows_kernel<CT, 16><<<n, 64, 0, s>>>(g, Nq, tstart, active, si0, rec, grid); break;
    }
}

// ----------------------------------------------------------------------------
// Pencil transposes between the 1D transforms. All four move 32×32 (or
// 32×16-pair) tiles through LDS so that both the loads and the stores run
// along contiguous rows.

// Forward, after the x-pass: Z = FFT_x(row 2p + i·row 2p+1) for every pair p.
// Writes ypen[c][kx][z][y] (y contiguous) for the retained kx ∈ [0, nkx):
// A_k = (Z_k + conj Z_{-k})/2 for y = 2p, B_k = (Z_k - conj Z_{-k})/(2i) for y = 2p+1.
// Grid: (⌈nkx/32⌉, G1/32, 3·G2), 256 threads.
template <typename CT>
__global__ __launch_bounds__(256) void pair_split_to_ypencils(const Cx<CT> *__restrict__ zrow,
                                                              Cx<CT> *__restrict__ ypen, int G0, int G1, int G2,
                                                              int nkx) {
    __shared__ Cx<CT> sk[16][33], sm[16][33];
    const int kx0 = blockIdx.x * 32, p0 = blockIdx.y * 16;
    const int cz = blockIdx.z, c = cz / G2, z = cz - c * G2;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const Cx<CT> *rows = zrow + (int64_t)cz * (G1 >> 1) * G0;
    const int kx = kx0 + tx;
    if (kx < nkx) {
        const int km = kx == 0 ? 0 : G0 - kx;
        for (int pp = ty; pp < 16; pp += 8) {
            const Cx<CT> *r = rows + (int64_t)(p0 + pp) * G0;
            sk[pp][tx] = r[kx];
            sm[pp][tx] = r[km];
        }
    }
    __syncthreads();
    const int pp = tx >> 1, par = tx & 1;
    for (int kk = ty; kk < 32 && kx0 + kk < nkx; kk += 8) {
        const Cx<CT> a = sk[pp][kk], b = sm[pp][kk];
        Cx<CT> o;
        if (par == 0) {
            o.re = (a.re + b.re) * (CT)0.5;
            o.im = (a.im - b.im) * (CT)0.5;
        } else {
            o.re = (a.im + b.im) * (CT)0.5;
            o.im = (b.re - a.re) * (CT)0.5;
        }
        ypen[(((int64_t)c * nkx + kx0 + kk) * G2 + z) * G1 + 2 * p0 + tx] = o;
    }
}

// Forward, after the y-pass: ypen[plane][z][iy] -> zpen[plane][j][z] for the
// retained ky = j (j ≤ K1) or j - nky (j > K1), plane = c·nkx + kx.
// Grid: (G2/32, ⌈nky/32⌉, 3·nkx).
template <typename CT>
__global__ __launch_bounds__(256) void ypencils_to_zpencils(const Cx<CT> *__restrict__ ypen,
                                                            Cx<CT> *__restrict__ zpen, int G1, int G2, int nky,
                                                            int K1) {
    __shared__ Cx<CT> s[32][33];
    const int z0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int64_t plane = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const Cx<CT> *src = ypen + plane * G2 * G1;
    Cx<CT> *dst = zpen + plane * nky * G2;
    const int j = j0 + tx;
    if (j < nky) {
        const int iy = j <= K1 ? j : G1 + j - nky;
        for (int zz = ty; zz < 32; zz += 8) s[zz][tx] = src[(int64_t)(z0 + zz) * G1 + iy];
    }
    __syncthreads();
    for (int jj = ty; jj < 32 && j0 + jj < nky; jj += 8) dst[(int64_t)(j0 + jj) * G2 + z0 + tx] = s[tx][jj];
}

// Inverse, after the z-pass: zpen[plane][j][z] -> ypen[plane][z][iy] for every
// iy, zero where ky is not retained. Grid: (G2/32, G1/32, 3·nkx).
template <typename CT>
__global__ __launch_bounds__(256) void zpencils_to_ypencils(const Cx<CT> *__restrict__ zpen,
                                                            Cx<CT> *__restrict__ ypen, int G1, int G2, int nky,
                                                            int K1) {
    __shared__ Cx<CT> s[32][33];
    const int z0 = blockIdx.x * 32, iy0 = blockIdx.y * 32;
    const int64_t plane = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const Cx<CT> *src = zpen + plane * nky * G2;
    Cx<CT> *dst = ypen + plane * G2 * G1;
    for (int jj = ty; jj < 32; jj += 8) {
        const int iy = iy0 + jj;
        const int ky = iy <= (G1 >> 1) ? iy : iy - G1;
        Cx<CT> v{(CT)0, (CT)0};
        if (ky >= -K1 && ky <= K1) v = src[(int64_t)(ky >= 0 ? ky : ky + nky) * G2 + z0 + tx];
        s[jj][tx] = v;
    }
    __syncthreads();
    for (int zz = ty; zz < 32; zz += 8) dst[(int64_t)(z0 + zz) * G1 + iy0 + tx] = s[tx][zz];
}

// Inverse, after the y-pass: rebuild each pair row Z_k = A_k + i·B_k for every
// k ∈ [0, G0): A, B from ypen for k < nkx, their conjugates at G0 - k for
// k > G0 - nkx (Hermitian symmetry of real rows), zero otherwise; the DC term
// keeps only its real part (what a c2r transform uses).
// Grid: (G0/32, G1/32, 3·G2).
template <typename CT>
__global__ __launch_bounds__(256) void ypencils_to_pair_rows(const Cx<CT> *__restrict__ ypen,
                                                             Cx<CT> *__restrict__ zrow, int G0, int G1, int G2,
                                                             int nkx) {
    __shared__ Cx<CT> s[32][33];
    const int k0 = blockIdx.x * 32, p0 = blockIdx.y * 16;
    const int cz = blockIdx.z, c = cz / G2, z = cz - c * G2;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int kk = ty; kk < 32; kk += 8) {
        const int k = k0 + kk;
        int kx = -1;
        bool cj = false;
        if (k < nkx) {
            kx = k;
        } else if (k > G0 - nkx) {
            kx = G0 - k;
            cj = true;
        }
        Cx<CT> v{(CT)0, (CT)0};
        if (kx >= 0) {
            v = ypen[(((int64_t)c * nkx + kx) * G2 + z) * G1 + 2 * p0 + tx];
            if (cj) v.im = -v.im;
            if (k == 0) v.im = (CT)0;
        }
        s[kk][tx] = v;
    }
    __syncthreads();
    Cx<CT> *rows = zrow + (int64_t)cz * (G1 >> 1) * G0;
    for (int pp = ty; pp < 16; pp += 8) {
        const Cx<CT> A = s[tx][2 * pp], B = s[tx][2 * pp + 1];
        rows[(int64_t)(p0 + pp) * G0 + k0 + tx] = Cx<CT>{A.re - B.im, A.im + B.re};
    }
}

// ----------------------------------------------------------------------------
// Fused Fourier-space kernel, in place on the z-pencils [3][nkx][nky][G2].

template <typename CT>
struct FourierArgs {
    int G[3];
    int nkx, nky;
    int64_t cstride;  // complex values per component
    int kmax[3];      // retained: kx ∈ [0, kmax0], |ky| ≤ kmax1, |kz| ≤ kmax2
    const double *dx, *dy, *dz;  // 1/ψ̂ per retained mode (dy, dz offset by kmax)
    double kscale[3];            // 2π/L
    double pref;                 // Γ/V
    double inv4a2;               // 1/(4α²)
    double kmax2_trunc;          // < 0: no spherical truncation
    int do_vel, do_psi;
    // unsmoothed field kept for the Fourier-space consumers (velocity if do_vel, else streamfunction):
    // uhat[3][nky][nkz][nkx], kz index jz ≤ kmax2 → kz = jz, else jz - nkz
    CT *uhat;
    int keep_uhat, nkz;
    int64_t ucstride;
};

template <typename CT>
__global__ void fourier_kernel(FourierArgs<CT> a, CT *__restrict__ spec, CT *__restrict__ spec2) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.cstride) return;
    const int iz = (int)(idx % a.G[2]);
    const int j = (int)((idx / a.G[2]) % a.nky);
    const int kx = (int)(idx / ((int64_t)a.G[2] * a.nky));
    const int ky = j <= a.kmax[1] ? j : j - a.nky;
    const int kz = iz <= a.G[2] / 2 ? iz : iz - a.G[2];
    bool keep = kz >= -a.kmax[2] && kz <= a.kmax[2];
    const double KX = a.kscale[0] * kx, KY = a.kscale[1] * ky, KZ = a.kscale[2] * kz;
    const double k2 = KX * KX + KY * KY + KZ * KZ;
    if (a.kmax2_trunc >= 0 && k2 > a.kmax2_trunc) keep = false;
    CT *c1 = spec + 2 * idx;   // complex (re, im) of component 0
    CT *c2 = spec2 + 2 * idx;
    const int64_t cs = 2 * a.cstride;
    // slot of this mode in the kept field (the retained kz only)
    CT *cu = nullptr;
    if (a.keep_uhat && kz >= -a.kmax[2] && kz <= a.kmax[2]) {
        const int jz = kz >= 0 ? kz : kz + a.nkz;
        cu = a.uhat + 2 * (((int64_t)j * a.nkz + jz) * a.nkx + kx);
    }
    if (!keep || k2 == 0) {
        if (cu) {
#pragma unroll
            for (int d = 0; d < 3; d++) cu[2 * d * a.ucstride] = cu[2 * d * a.ucstride + 1] = (CT)0;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (a.do_vel) {
                c1[d * cs] = (CT)0;
                c1[d * cs + 1] = (CT)0;
            }
            if (a.do_psi) {
                c2[d * cs] = (CT)0;
                c2[d * cs + 1] = (CT)0;
            }
        }
        // streamfunction-only: the vorticity buffer is not an output, leave it
        return;
    }
    const double D = a.dx[kx] * a.dy[ky + a.kmax[1]] * a.dz[kz + a.kmax[2]];
    // ω̂ = Γ/V · B̂ / ψ̂ ; output = op(ω̂) · e^{-k²/4α²} / ψ̂
    const double s = a.pref * D * D * exp(-k2 * a.inv4a2) / k2;
    double wr[3], wi[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        wr[d] = (double)c1[d * cs] * s;
        wi[d] = (double)c1[d * cs + 1] * s;
    }
    if (cu) {
        // unsmoothed field: ω̂/k² (one deconvolution, no Ewald factor), then i k × for the velocity
        const double su = a.pref * D / k2;
        double ur[3], ui[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            ur[d] = (double)c1[d * cs] * su;
            ui[d] = (double)c1[d * cs + 1] * su;
        }
        if (a.do_vel) {
            const double ax = KY * ur[2] - KZ * ur[1], bx = KY * ui[2] - KZ * ui[1];
            const double ay = KZ * ur[0] - KX * ur[2], by = KZ * ui[0] - KX * ui[2];
            const double az = KX * ur[1] - KY * ur[0], bz = KX * ui[1] - KY * ui[0];
            ur[0] = -bx, ui[0] = ax, ur[1] = -by, ui[1] = ay, ur[2] = -bz, ui[2] = az;
        }
        const int64_t us = 2 * a.ucstride;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            cu[d * us] = (CT)ur[d];
            cu[d * us + 1] = (CT)ui[d];
        }
    }
    if (a.do_psi) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            c2[d * cs] = (CT)wr[d];
            c2[d * cs + 1] = (CT)wi[d];
        }
    }
    if (a.do_vel) {
        // i k × w: (i k × w)_x = i (KY w_z - KZ w_y), i (a + ib) = -b + ia
        const double ax = KY * wr[2] - KZ * wr[1], bx = KY * wi[2] - KZ * wi[1];
        const double ay = KZ * wr[0] - KX * wr[2], by = KZ * wi[0] - KX * wi[2];
        const double az = KX * wr[1] - KY * wr[0], bz = KX * wi[1] - KY * wi[0];
        c1[0] = (CT)(-bx);
        c1[1] = (CT)ax;
        c1[cs] = (CT)(-by);
        c1[cs + 1] = (CT)ay;
        c1[2 * cs] = (CT)(-bz);
        c1[2 * cs + 1] = (CT)az;
    }
}

// ----------------------------------------------------------------------------
// Consumers of the kept Fourier field.

// z-pencils of the band-filtered field, ready for the inverse transforms: the
// kept û on kmin² ≤ k² ≤ kmax² times the type-2 deconvolution, zero elsewhere.
template <typename CT>
__global__ void band_to_zpencils(FourierArgs<CT> a, double kmin2, double kmax2, double cf, double cc,
                                 CT *__restrict__ zp) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.cstride) return;
    const int iz = (int)(idx % a.G[2]);
    const int j = (int)((idx / a.G[2]) % a.nky);
    const int kx = (int)(idx / ((int64_t)a.G[2] * a.nky));
    const int ky = j <= a.kmax[1] ? j : j - a.nky;
    const int kz = iz <= a.G[2] / 2 ? iz : iz - a.G[2];
    const double KX = a.kscale[0] * kx, KY = a.kscale[1] * ky, KZ = a.kscale[2] * kz;
    const double k2 = KX * KX + KY * KY + KZ * KZ;
    const bool in = kz >= -a.kmax[2] && kz <= a.kmax[2] && k2 >= kmin2 && k2 <= kmax2;
    const int64_t cs = 2 * a.cstride, us = 2 * a.ucstride;
    CT *o = zp + 2 * idx;
    if (!in) {
#pragma unroll
        for (int d = 0; d < 3; d++) o[d * cs] = o[d * cs + 1] = (CT)0;
        return;
    }
    double D = a.dx[kx] * a.dy[ky + a.kmax[1]] * a.dz[kz + a.kmax[2]];
    if (cf > 0) D *= exp(-cf * k2) - exp(-cc * k2);  // difference of two Ewald Gaussians
    const int jz = kz >= 0 ? kz : kz + a.nkz;
    const CT *u = a.uhat + 2 * (((int64_t)j * a.nkz + jz) * a.nkx + kx);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        o[d * cs] = (CT)((double)u[d * us] * D);
        o[d * cs + 1] = (CT)((double)u[d * us + 1] * D);
    }
}

// Energy spectrum of the kept field (Diagnostics/spectra.jl:89-196): mode k goes
// to bin n = trunc(|k|/Δk + 1/2) with weight f·|û|²/Δk (f = 1/2 on the kx = 0
// plane, 1 elsewhere; from the streamfunction |k × ψ̂|² = k²|ψ̂|²). One partial
// spectrum per block (LDS bins), summed over blocks in a fixed order.
template <typename CT>
__global__ void spectrum_partial(const CT *__restrict__ uhat, int64_t ucs, int nkx, int nkz, int K1, int K2,
                                 double kx_scale, double ky_scale, double kz_scale, double dk_inv, int nk,
                                 int from_psi, double *__restrict__ partial) {
    extern __shared__ double bins[];
    for (int b = threadIdx.x; b < nk; b += blockDim.x) bins[b] = 0.0;
    __syncthreads();
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < ucs;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(idx % nkx);
        const int jz = (int)((idx / nkx) % nkz);
        const int j = (int)(idx / ((int64_t)nkx * nkz));
        const int ky = j <= K1 ? j : j - (2 * K1 + 1);
        const int kz = jz <= K2 ? jz : jz - nkz;
        const double KX = kx_scale * kx, KY = ky_scale * ky, KZ = kz_scale * kz;
        const double k2 = KX * KX + KY * KY + KZ * KZ;
        const int n = (int)(sqrt(k2) * dk_inv + 0.5);
        if (n >= nk) continue;
        double u2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double re = uhat[2 * (d * ucs + idx)], im = uhat[2 * (d * ucs + idx) + 1];
            u2 += re * re + im * im;
        }
        if (from_psi) u2 *= k2;
        atomicAdd(&bins[n], (kx == 0 ? 0.5 : 1.0) * u2 * dk_inv);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nk; b += blockDim.x) partial[(int64_t)blockIdx.x * nk + b] = bins[b];
}

__global__ void spectrum_reduce(const double *__restrict__ partial, int nblocks, int nk, double *__restrict__ Ek) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nk) return;
    double s = 0.0;
    for (int k = 0; k < nblocks; k++) s += partial[(int64_t)k * nk + b];
    Ek[b] = s;
}

// ----------------------------------------------------------------------------
// Interpolation (type 2).

// Type-2 interpolation with a W³ stencil (W = 2m): one wavefront (block) per node.
// Lanes 0..3W-1 evaluate the node's 3·W kernel weights once and share them through
// LDS (exact copies); lane (a, b) = (lane % W, lane / W) then owns stencil column
// x = i0x + a and rows y = i0y + b + R·pass (R = ⌊64/W⌋ rows per pass) and walks the
// W z-planes, so a load instruction touches R rows of W consecutive values; the
// partial sums are combined by DPP steps in a fixed order (total in lane 63).
template <typename CT, int W>
__global__ __launch_bounds__(64) void interp_kernel_wave(LRGeom g, int64_t Np, const double *__restrict__ nodes,
                                                         CT beta, const CT *__restrict__ grid,
                                                         double *__restrict__ out) {
    constexpr int M = W / 2;
    constexpr int R = 64 / W;                  // y rows per pass
    constexpr int PASSES = (W + R - 1) / R;
    __shared__ CT s_w[3][W];
    const int lane = threadIdx.x;
    const int a = lane % W, b = lane / W;
    const bool lane_on = b < R;                // lanes beyond R·W idle
    // Blocks are dealt to the 8 XCDs round-robin; remap so that each XCD works
    // on a contiguous range of nodes (consecutive nodes share stencil rows,
    // which then stay in that XCD's L2).
    const int64_t per = ((int64_t)gridDim.x + 7) / 8;
    const int64_t i = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (i >= Np) return;  // the whole (one-wavefront) block
    const double ux = nodes[i] * g.scale[0];
    const double uy = nodes[Np + i] * g.scale[1];
    const double uz = nodes[2 * Np + i] * g.scale[2];
    const int i0x = (int)floor(ux) - M + 1, i0y = (int)floor(uy) - M + 1, i0z = (int)floor(uz) - M + 1;
    if (lane < 3 * W) {
        const int t = lane / W, c = lane - t * W;
        const double u = t == 0 ? ux : t == 1 ? uy : uz;
        const int i0 = t == 0 ? i0x : t == 1 ? i0y : i0z;
        s_w[t][c] = es_kernel<CT>((CT)(((double)(i0 + c) - u) / (double)M), beta);
    }
    __syncthreads();
    const CT wx = s_w[0][a];
    CT wz[W];
    int64_t gzo[W];
    const int64_t plane = (int64_t)g.G[0] * g.G[1];  // values per z-plane
#pragma unroll
    for (int c = 0; c < W; c++) {
        wz[c] = s_w[2][c];
        gzo[c] = (int64_t)wrapi(i0z + c, g.G[2]) * plane;
    }
    const int64_t gx = 2 * (int64_t)wrapi(i0x + a, g.G[0]);
    CT acc0 = 0, acc1 = 0, acc2 = 0;
#pragma unroll
    for (int pass = 0; pass < PASSES; pass++) {
        const int bb = b + R * pass;
        if (lane_on && bb < W) {
            const CT wy = s_w[1][bb];
            const int iy = wrapi(i0y + bb, g.G[1]);
            const int64_t gy = (int64_t)(iy >> 1) * 2 * g.G[0] + (iy & 1);
            const CT wxy = wx * wy;
#pragma unroll
            for (int c = 0; c < W; c++) {
                const int64_t off = gzo[c] + gy + gx;
                const CT w = wxy * wz[c];
                acc0 += w * grid[off];
                acc1 += w * grid[g.cstride + off];
                acc2 += w * grid[2 * g.cstride + off];
            }
        }
    }
    acc0 = wave_sum_last(acc0);
    acc1 = wave_sum_last(acc1);
    acc2 = wave_sum_last(acc2);
    if (lane == 63) {
        out[i] = (double)acc0;
        out[Np + i] = (double)acc1;
        out[2 * Np + i] = (double)acc2;
    }
}

template <typename CT>
void launch_interp(const LRGeom &g, int64_t Np, const double *nodes, CT beta, const CT *grid, double *out,
                   hipStream_t s) {
    const unsigned nb = (unsigned)(8 * ((Np + 7) / 8));  // one wavefront per node and block
    switch (g.M) {
        case 2: interp_kernel_wave<CT, 4><<<nb, 64, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        case 3: interp_kernel_wave<CT, 6><<<nb, 64, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        case 4: interp_kernel_wave<CT, 8><<<nb, 64, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        case 5: interp_kernel_wave<CT, 10><<<nb, 64, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        case 6: interp_kernel_wave<CT, 12><<<nb, 64, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        case 7: interp_kernel_wave<CT, 14><<<nb, 64, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        default: interp_kernel_wave<CT, 16><<<nb, 64, 0, s>>>(g, Np, nodes, beta, grid, out); break;
    }
}

__global__ void add3(int64_t n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ o) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = a[i] + b[i];
}

// ----------------------------------------------------------------------------
// Host helpers

// Gauss–Legendre nodes/weights on [0, 1] by Newton iteration on P_n.
void gauss_legendre01(int n, std::vector<double> &x, std::vector<double> &w) {
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; i++) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double pp = 0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = (1.0 - z) / 2.0;  // map to [0, 1]
        w[i] = 1.0 / ((1.0 - z * z) * pp * pp);  // = (2/((1-z²)P'²))/2
    }
}

// ψ̂(k) = 2m ∫_0^1 φ(z) cos(2πkmz/G) dz
double es_fourier(int k, int G, int m, double beta, const std::vector<double> &x, const std::vector<double> &w) {
    double s = 0;
    for (size_t q = 0; q < x.size(); q++) {
        double z = x[q];
        double phi = std::exp(beta * (std::sqrt(1 - z * z) - 1));
        s += w[q] * phi * std::cos(2 * M_PI * k * m * z / G);
    }
    return 2 * m * s;
}

template <typename CT>
FourierArgs<CT> fourier_args(const vpbs_ctx *c) {
    FourierArgs<CT> fa;
    for (int d = 0; d < 3; d++) {
        fa.G[d] = c->G[d];
        fa.kmax[d] = c->kmax_mode[d];
        fa.kscale[d] = 2 * M_PI / c->p.Ls[d];
    }
    fa.nkx = c->nkx;
    fa.nky = c->nky;
    fa.cstride = c->zpen_stride;
    fa.dx = c->deconv;
    fa.dy = c->deconv + (c->kmax_mode[0] + 1);
    fa.dz = fa.dy + (2 * c->kmax_mode[1] + 1);
    fa.pref = c->p.Gamma / (c->p.Ls[0] * c->p.Ls[1] * c->p.Ls[2]);
    fa.inv4a2 = 1.0 / (4 * c->p.alpha * c->p.alpha);
    fa.kmax2_trunc = -1;
    if (c->p.truncate_spherical) {
        double km = 1e300;
        for (int d = 0; d < 3; d++) km = std::fmin(km, 2 * M_PI * ((c->p.Ns[d] - 1) / 2) / c->p.Ls[d]);
        fa.kmax2_trunc = km * km;  // params.jl:161-166
    }
    fa.do_vel = fa.do_psi = 0;
    fa.uhat = (CT *)c->uhat;
    fa.keep_uhat = 0;
    fa.nkz = c->nkz;
    fa.ucstride = c->uhat_stride;
    return fa;
}

LRGeom lr_geom(const vpbs_ctx *c) {
    LRGeom g;
    for (int d = 0; d < 3; d++) {
        g.G[d] = c->G[d];
        g.scale[d] = (double)c->G[d] / c->p.Ls[d];
        g.ntile[d] = c->ntile[d];
    }
    g.M = c->p.nufft_m;
    g.tx = c->p.precision == VPBS_F32 ? tile_x<float>() : tile_x<double>();
    g.cstride = c->comp_stride;
    return g;
}

// Inverse transforms of z-pencils zp into the real grid (z, y, x passes with the
// pencil transposes between them).
template <typename CT>
int inverse_to_grid(vpbs_ctx *c, Cx<CT> *zp) {
    if (c->fft_fused) return spectral_inverse(c, zp);
    hipStream_t s = c->s_lr;
    const int G0 = c->G[0], G1 = c->G[1], G2 = c->G[2];
    const int nkx = c->nkx, nky = c->nky, K1 = c->kmax_mode[1];
    Cx<CT> *rows = (Cx<CT> *)c->grid;
    Cx<CT> *ypen = (Cx<CT> *)c->ypen;
    auto fft = [&](int dim, void *buf) -> int {
        void *io[1] = {buf};
        if (rocfft_execute(c->plan[dim][1], io, nullptr, c->fft_info) != rocfft_status_success)
            return fail(c, VPBS_ERR_FFT, "rocfft_execute failed (dimension %d, inverse)", dim);
        return VPBS_OK;
    };
    VPBS_TRY(fft(2, zp));
    zpencils_to_ypencils<CT><<<dim3(G2 / 32, G1 / 32, 3 * nkx), 256, 0, s>>>(zp, ypen, G1, G2, nky, K1);
    VPBS_TRY(fft(1, ypen));
    ypencils_to_pair_rows<CT><<<dim3(G0 / 32, G1 / 32, 3 * G2), 256, 0, s>>>(ypen, rows, G0, G1, G2, nkx);
    VPBS_TRY(fft(0, rows));
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

// Node buffers of the slab interpolation, grown with Np.
int slab_node_buffers(vpbs_ctx *c) {
    const int64_t Np = c->Np;
    if (Np <= c->cap_nd && c->nd_key) return VPBS_OK;
    void **bufs[] = {(void **)&c->nd_key, (void **)&c->nd_perm, (void **)&c->nd_i0, &c->nd_rec, &c->slab_part};
    for (void **b : bufs)
        if (*b) {
            VPBS_HIP(c, hipFree(*b));
            *b = nullptr;
        }
    const size_t es = c->p.precision == VPBS_F32 ? sizeof(float) : sizeof(double);
    const size_t rec = (size_t)slab_rec_len(2 * c->p.nufft_m, (int)es);
    VPBS_HIP(c, hipMalloc((void **)&c->nd_key, sizeof(int32_t) * (size_t)Np));
    VPBS_HIP(c, hipMalloc((void **)&c->nd_perm, sizeof(int32_t) * (size_t)Np));
    VPBS_HIP(c, hipMalloc((void **)&c->nd_i0, sizeof(int32_t) * 3 * (size_t)Np));
    VPBS_HIP(c, hipMalloc(&c->nd_rec, es * rec * (size_t)Np));
    VPBS_HIP(c, hipMalloc(&c->slab_part, es * 3 * (size_t)spectral_slab_slots(c) * (size_t)Np));
    c->cap_nd = Np;
    return VPBS_OK;
}

template <typename CT>
int launch_lr(vpbs_ctx *c, uint32_t flags) {
    hipStream_t s = c->s_lr;
    const int64_t Np = c->Np, Nq = c->Nq;
    const bool vel = flags & VPBS_VELOCITY, psi = flags & VPBS_STREAMFUNCTION;
    const LRGeom g = lr_geom(c);
    const CT beta = (CT)c->es_beta;

    stage_begin(c, ST_LRSORT, s);
    lr_tile_keys<<<nblocks(Nq, 256), 256, 0, s>>>(Nq, c->points, g, c->lr_key);
    VPBS_TRY(radix_sort_bins(c, c->sort_lr, c->lr_key, Nq, c->ntiles, c->lr_start, c->lr_perm, s));
    launch_gather<CT>(Nq, c->lr_perm, c->points, c->charges, g, beta, c->lr_i0, (CT *)c->lr_w, s);
    stage_end(c, ST_LRSORT, s);
    // The slab-sorted node records are read only by the interpolation at the end: they are
    // sorted and gathered on the short-range stream (before its own work), beside this
    // stream's spreading, and the interpolation waits for them.
    const bool nodes_aside = c->slab_interp && Np > 0;
    if (nodes_aside) {
        hipStream_t s2 = c->s_sr;
        VPBS_TRY(slab_node_buffers(c));
        stage_begin(c, ST_NDSORT, s2);
        slab_keys<<<nblocks(Np, 256), 256, 0, s2>>>(Np, c->nodes, g, c->nd_key);
        VPBS_TRY(radix_sort_bins(c, c->sort_nd, c->nd_key, Np, c->nslabs, c->nd_start, c->nd_perm, s2));
        launch_gather_nodes<CT>(Np, c->nd_perm, c->nodes, g, beta, c->nd_i0, (CT *)c->nd_rec, s2);
        stage_end(c, ST_NDSORT, s2);
        VPBS_HIP(c, hipGetLastError());
        VPBS_HIP(c, hipEventRecord(c->ev_nodes, s2));
    }

    CT *grid = (CT *)c->grid;
    stage_begin(c, ST_SPREAD, s);
    tile_active<<<nblocks(c->ntiles, 256), 256, 0, s>>>(g, c->ntiles, c->lr_start, c->lr_active);
    const int64_t nb_spread = ((c->ntiles + 7) / 8) * 8;  // xcd_tile() maps 8·⌈n/8⌉ blocks onto the tiles
    launch_spread<CT>(g, nb_spread, Nq, c->lr_start, c->lr_active, c->lr_i0, (const CT *)c->lr_w, grid, s);
    stage_end(c, ST_SPREAD, s);
    VPBS_HIP(c, hipGetLastError());

    if (c->fft_fused) {
        stage_begin(c, ST_FFTF, s);
        VPBS_TRY(spectral_forward(c));
        stage_end(c, ST_FFTF, s);
        stage_begin(c, ST_FOURIER, s);
        VPBS_TRY(spectral_fourier(c, flags));
        stage_end(c, ST_FOURIER, s);
        // inverse transforms to the nodes: with slab interpolation the inverse x transform
        // runs inside the interpolation kernel (the real grid is not written)
        stage_begin(c, ST_FFTB, s);
        if (vel) VPBS_TRY(c->slab_interp ? spectral_y_inverse(c, c->zpen) : spectral_inverse(c, c->zpen));
        stage_end(c, ST_FFTB, s);
        if (nodes_aside) VPBS_HIP(c, hipStreamWaitEvent(s, c->ev_nodes, 0));
        stage_begin(c, ST_INTERP, s);
        if (vel) {
            if (c->slab_interp)
                VPBS_TRY(spectral_x_inverse_interp(c, c->v_lr));
            else
                launch_interp<CT>(g, Np, c->nodes, beta, grid, c->v_lr, s);
        }
        if (psi) {
            if (c->slab_interp) {
                VPBS_TRY(spectral_y_inverse(c, c->zpen2));
                VPBS_TRY(spectral_x_inverse_interp(c, c->psi_lr));
            } else {
                VPBS_TRY(spectral_inverse(c, c->zpen2));
                launch_interp<CT>(g, Np, c->nodes, beta, grid, c->psi_lr, s);
            }
        }
        stage_end(c, ST_INTERP, s);
        VPBS_HIP(c, hipGetLastError());
        return VPBS_OK;
    }

    const int G0 = c->G[0], G1 = c->G[1], G2 = c->G[2];
    const int nkx = c->nkx, nky = c->nky, K1 = c->kmax_mode[1];
    Cx<CT> *rows = (Cx<CT> *)c->grid;
    Cx<CT> *ypen = (Cx<CT> *)c->ypen;
    auto fft = [&](int dim, int dir, void *buf) -> int {
        void *io[1] = {buf};
        if (rocfft_execute(c->plan[dim][dir], io, nullptr, c->fft_info) != rocfft_status_success)
            return fail(c, VPBS_ERR_FFT, "rocfft_execute failed (dimension %d, %s)", dim,
                        dir ? "inverse" : "forward");
        return VPBS_OK;
    };
    auto inverse = [&](Cx<CT> *zp) -> int { return inverse_to_grid<CT>(c, zp); };

    stage_begin(c, ST_FFTF, s);
    Cx<CT> *zpen = (Cx<CT> *)c->zpen;
    VPBS_TRY(fft(0, 0, rows));
    pair_split_to_ypencils<CT><<<dim3((nkx + 31) / 32, G1 / 32, 3 * G2), 256, 0, s>>>(rows, ypen, G0, G1, G2, nkx);
    VPBS_TRY(fft(1, 0, ypen));
    ypencils_to_zpencils<CT><<<dim3(G2 / 32, (nky + 31) / 32, 3 * nkx), 256, 0, s>>>(ypen, zpen, G1, G2, nky, K1);
    VPBS_TRY(fft(2, 0, zpen));
    stage_end(c, ST_FFTF, s);
    VPBS_HIP(c, hipGetLastError());

    FourierArgs<CT> fa = fourier_args<CT>(c);
    fa.do_vel = vel;
    fa.do_psi = psi;
    fa.keep_uhat = (flags & VPBS_KEEP_UHAT) ? 1 : 0;
    Cx<CT> *zpen2 = psi ? (Cx<CT> *)c->zpen2 : zpen;
    stage_begin(c, ST_FOURIER, s);
    fourier_kernel<CT><<<nblocks(fa.cstride, 256), 256, 0, s>>>(fa, (CT *)zpen, (CT *)zpen2);
    stage_end(c, ST_FOURIER, s);
    VPBS_HIP(c, hipGetLastError());

    // velocity: z-pencils -> real grid -> nodes; then the streamfunction reuses the grid
    stage_begin(c, ST_FFTB, s);
    if (vel) VPBS_TRY(inverse(zpen));
    stage_end(c, ST_FFTB, s);
    stage_begin(c, ST_INTERP, s);
    if (vel) launch_interp<CT>(g, Np, c->nodes, beta, grid, c->v_lr, s);
    if (psi) {
        VPBS_TRY(inverse(zpen2));
        launch_interp<CT>(g, Np, c->nodes, beta, grid, c->psi_lr, s);
    }
    stage_end(c, ST_INTERP, s);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

bool rocfft_ok(rocfft_status st) { return st == rocfft_status_success; }

}  // namespace

double es_beta(int m, double sigma) { return 0.98 * M_PI * m * (2.0 - 1.0 / sigma); }

int oversampled_size(int N, double sigma) {
    // smallest multiple of 32 with no prime factor above 5, ≥ ⌈σN⌉
    int n = (int)std::ceil(sigma * N - 1e-9);
    for (;; n++) {
        if (n % 32) continue;
        int r = n;
        for (int p : {2, 3, 5})
            while (r % p == 0) r /= p;
        if (r == 1) return n;
    }
}

// Device address of `scratch` on the context's GPU (current here), taken once per device.
int scratch_address(vpbs_ctx *c) {
    if (c->device < 0 || c->device >= MAX_DEVICES)
        return fail(c, VPBS_ERR_INVALID, "device %d: at most %d devices", c->device, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        void *ptr = nullptr;
        VPBS_HIP(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[c->device] = (double *)ptr;
    }
    c->spec_scratch = g_scratch_ptr[c->device];
    return VPBS_OK;
}

int longrange_setup(vpbs_ctx *c) {
    const int m = c->p.nufft_m;
    VPBS_TRY(scratch_address(c));
    for (int d = 0; d < 3; d++) {
        c->G[d] = oversampled_size(c->p.Ns[d], c->p.nufft_sigma);
        if (c->G[d] < 4 * m) c->G[d] = ((4 * m + 31) / 32) * 32;
        const int tx = c->p.precision == VPBS_F32 ? tile_x<float>() : tile_x<double>();
        c->ntile[d] = c->G[d] / (d == 0 ? tx : (d == 1 ? TY : TZ));
    }
    c->ntiles = (int64_t)c->ntile[0] * c->ntile[1] * c->ntile[2];
    c->nslabs = (int64_t)(c->G[1] / SLAB_HY) * (c->G[2] / SLAB_HZ);
    c->comp_stride = (int64_t)c->G[0] * c->G[1] * c->G[2];
    c->es_beta = es_beta(m, c->p.nufft_sigma);
    // retained modes (longrange.jl:556-583)
    c->kmax_mode[0] = c->p.Ns[0] / 2 - 1;
    for (int d = 1; d < 3; d++)
        c->kmax_mode[d] = (c->p.Ns[d] % 2 == 0) ? c->p.Ns[d] / 2 - 1 : (c->p.Ns[d] - 1) / 2;
    c->nkx = c->kmax_mode[0] + 1;
    c->nky = 2 * c->kmax_mode[1] + 1;
    c->ypen_stride = (int64_t)c->nkx * c->G[2] * c->G[1];
    c->zpen_stride = (int64_t)c->nkx * c->nky * c->G[2];
    c->nkz = 2 * c->kmax_mode[2] + 1;
    c->uhat_stride = (int64_t)c->nkx * c->nky * c->nkz;

    // deconvolution tables 1/ψ̂(k)
    std::vector<double> gx, gw;
    gauss_legendre01(128, gx, gw);
    std::vector<double> tab;
    for (int k = 0; k <= c->kmax_mode[0]; k++) tab.push_back(1.0 / es_fourier(k, c->G[0], m, c->es_beta, gx, gw));
    for (int d = 1; d < 3; d++)
        for (int k = -c->kmax_mode[d]; k <= c->kmax_mode[d]; k++)
            tab.push_back(1.0 / es_fourier(k, c->G[d], m, c->es_beta, gx, gw));
    VPBS_HIP(c, hipMalloc((void **)&c->deconv, sizeof(double) * tab.size()));
    VPBS_HIP(c, hipMemcpy(c->deconv, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));

    const size_t es = c->p.precision == VPBS_F32 ? sizeof(float) : sizeof(double);
    VPBS_HIP(c, hipMalloc(&c->grid, es * 3 * (size_t)c->comp_stride));
    VPBS_HIP(c, hipMalloc(&c->ypen, 2 * es * 3 * (size_t)c->ypen_stride));
    VPBS_HIP(c, hipMalloc(&c->zpen, 2 * es * 3 * (size_t)c->zpen_stride));

    c->fft_fused = (c->p.fft == VPBS_FFT_FUSED || c->p.fft == VPBS_FFT_SEPARATE) && spectral_supported(c);
    c->slab_interp = false;
    if (c->fft_fused) {
        VPBS_TRY(spectral_setup(c));
        c->slab_interp = spectral_slab_supported(c);
        c->plans_ready = true;
        return VPBS_OK;
    }

    // rocFFT: in-place batched contiguous c2c along each direction
    const rocfft_precision prec = c->p.precision == VPBS_F32 ? rocfft_precision_single : rocfft_precision_double;
    const size_t len[3] = {(size_t)c->G[0], (size_t)c->G[1], (size_t)c->G[2]};
    const size_t batch[3] = {3 * (size_t)c->G[2] * (c->G[1] / 2), 3 * (size_t)c->nkx * c->G[2],
                             3 * (size_t)c->nkx * c->nky};
    c->fft_work_size = 0;
    for (int d = 0; d < 3; d++) {
        for (int dir = 0; dir < 2; dir++) {
            rocfft_plan_description desc = nullptr;
            if (!rocfft_ok(rocfft_plan_description_create(&desc)))
                return fail(c, VPBS_ERR_FFT, "rocfft_plan_description_create failed");
            const size_t stride[1] = {1};
            const bool ok =
                rocfft_ok(rocfft_plan_description_set_data_layout(
                    desc, rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved, nullptr,
                    nullptr, 1, stride, len[d], 1, stride, len[d])) &&
                rocfft_ok(rocfft_plan_create(&c->plan[d][dir], rocfft_placement_inplace,
                                             dir ? rocfft_transform_type_complex_inverse
                                                 : rocfft_transform_type_complex_forward,
                                             prec, 1, &len[d], batch[d], desc));
            rocfft_plan_description_destroy(desc);
            if (!ok) return fail(c, VPBS_ERR_FFT, "rocfft_plan_create failed (length %zu, batch %zu)", len[d], batch[d]);
            size_t w = 0;
            rocfft_plan_get_work_buffer_size(c->plan[d][dir], &w);
            c->fft_work_size = std::max(c->fft_work_size, w);
        }
    }
    if (c->fft_work_size) VPBS_HIP(c, hipMalloc(&c->fft_work, c->fft_work_size));
    if (!rocfft_ok(rocfft_execution_info_create(&c->fft_info)))
        return fail(c, VPBS_ERR_FFT, "rocfft_execution_info_create failed");
    rocfft_execution_info_set_stream(c->fft_info, c->s_lr);
    if (c->fft_work_size) rocfft_execution_info_set_work_buffer(c->fft_info, c->fft_work, c->fft_work_size);
    c->plans_ready = true;
    return VPBS_OK;
}

int longrange_compute(vpbs_ctx *c, uint32_t flags) {
    const bool psi = flags & VPBS_STREAMFUNCTION;
    const size_t es = c->p.precision == VPBS_F32 ? sizeof(float) : sizeof(double);
    if (psi && !c->zpen2) VPBS_HIP(c, hipMalloc(&c->zpen2, 2 * es * 3 * (size_t)c->zpen_stride));
    const bool keep = flags & VPBS_KEEP_UHAT;
    if (keep && !c->uhat) VPBS_HIP(c, hipMalloc(&c->uhat, 2 * es * 3 * (size_t)c->uhat_stride));
    c->uhat_quantity = 0;
    const int rc = c->p.precision == VPBS_F32 ? launch_lr<float>(c, flags) : launch_lr<double>(c, flags);
    if (rc == VPBS_OK && keep) {
        c->uhat_quantity = (flags & VPBS_VELOCITY) ? VPBS_UHAT_VELOCITY : VPBS_UHAT_STREAMFUNCTION;
        c->uhat_version = c->points_version;
    }
    return rc;
}

namespace {

// Kept field on kmin² ≤ k² ≤ kmax² (weighted by e^{−cf·k²} − e^{−cc·k²} when cf > 0) on the
// real grid.
template <typename CT>
int band_grid(vpbs_ctx *c, double kmin2, double kmax2, double cf, double cc) {
    hipStream_t s = c->s_lr;
    if (c->fft_fused) {
        VPBS_TRY(spectral_band(c, kmin2, kmax2, cf, cc));
    } else {
        FourierArgs<CT> fa = fourier_args<CT>(c);
        band_to_zpencils<CT><<<nblocks(fa.cstride, 256), 256, 0, s>>>(fa, kmin2, kmax2, cf, cc, (CT *)c->zpen);
        VPBS_HIP(c, hipGetLastError());
        VPBS_TRY(inverse_to_grid<CT>(c, (Cx<CT> *)c->zpen));
    }
    return VPBS_OK;
}

template <typename CT>
int band_impl(vpbs_ctx *c, double kmin, double kmax, double *v_out) {
    hipStream_t s = c->s_lr;
    VPBS_TRY(band_grid<CT>(c, kmin * kmin, kmax * kmax, 0.0, 0.0));
    launch_interp<CT>(lr_geom(c), c->Np, c->nodes, (CT)c->es_beta, (const CT *)c->grid, c->v_lr, s);
    VPBS_HIP(c, hipGetLastError());
    VPBS_HIP(c, hipMemcpyAsync(v_out, c->v_lr, sizeof(double) * 3 * (size_t)c->Np, hipMemcpyDeviceToHost, s));
    VPBS_HIP(c, hipStreamSynchronize(s));
    return VPBS_OK;
}

// Partial spectra of nb blocks into partial [nb][nk], their sum into out [nk].
template <typename CT>
void launch_spectrum(vpbs_ctx *c, int nk, double dk, int nb, double *partial, double *out, hipStream_t s) {
    spectrum_partial<CT><<<nb, 256, sizeof(double) * nk, s>>>(
        (const CT *)c->uhat, c->uhat_stride, c->nkx, c->nkz, c->kmax_mode[1], c->kmax_mode[2], 2 * M_PI / c->p.Ls[0],
        2 * M_PI / c->p.Ls[1], 2 * M_PI / c->p.Ls[2], 1.0 / dk, nk, c->uhat_quantity == VPBS_UHAT_STREAMFUNCTION,
        partial);
    spectrum_reduce<<<nblocks(nk, 256), 256, 0, s>>>(partial, nb, nk, out);
}

template <typename CT>
int spectrum_impl(vpbs_ctx *c, int nk, double dk, double *Ek) {
    hipStream_t s = c->s_lr;
    const int nb = 1024;
    double *partial = nullptr, *out = nullptr;
    VPBS_HIP(c, hipMallocAsync((void **)&partial, sizeof(double) * ((size_t)nb * nk + nk), s));
    out = partial + (size_t)nb * nk;
    launch_spectrum<CT>(c, nk, dk, nb, partial, out, s);
    VPBS_HIP(c, hipGetLastError());
    VPBS_HIP(c, hipMemcpyAsync(Ek, out, sizeof(double) * nk, hipMemcpyDeviceToHost, s));
    VPBS_HIP(c, hipFreeAsync(partial, s));
    VPBS_HIP(c, hipStreamSynchronize(s));
    return VPBS_OK;
}

// The same with the partial spectra and the sum in the GPU's `scratch` array (nk ≤
// SPEC_NA_MAXBINS): no allocation. Every partial bin and every output bin is written by the
// kernels before it is read, so nothing left by an earlier call is used.
template <typename CT>
int spectrum_noalloc_impl(vpbs_ctx *c, int nk, double dk, double *Ek) {
    hipStream_t s = c->s_lr;
    double *partial = c->spec_scratch, *out = partial + (size_t)SPEC_NA_BLOCKS * nk;
    launch_spectrum<CT>(c, nk, dk, SPEC_NA_BLOCKS, partial, out, s);
    VPBS_HIP(c, hipGetLastError());
    VPBS_HIP(c, hipMemcpyAsync(Ek, out, sizeof(double) * nk, hipMemcpyDefault, s));
    VPBS_HIP(c, hipStreamSynchronize(s));
    return VPBS_OK;
}

template <typename CT>
int get_uhat_impl(vpbs_ctx *c, double *out) {
    const int64_t n = 3 * c->uhat_stride;
    std::vector<CT> h(2 * (size_t)n);
    VPBS_HIP(c, hipMemcpy(h.data(), c->uhat, sizeof(CT) * 2 * (size_t)n, hipMemcpyDeviceToHost));
    const int64_t Nkx = c->p.Ns[0] / 2 + 1, N1 = c->p.Ns[1], N2 = c->p.Ns[2];
    const int64_t K = Nkx * N1 * N2;
    std::fill(out, out + 2 * 3 * K, 0.0);
    const int K1 = c->kmax_mode[1], K2 = c->kmax_mode[2];
    for (int d = 0; d < 3; d++)
        for (int kx = 0; kx < c->nkx; kx++)
            for (int j = 0; j < c->nky; j++) {
                const int ky = j <= K1 ? j : j - c->nky;
                const int64_t iy = ky >= 0 ? ky : ky + N1;
                for (int jz = 0; jz < c->nkz; jz++) {
                    const int kz = jz <= K2 ? jz : jz - c->nkz;
                    const int64_t iz = kz >= 0 ? kz : kz + N2;
                    const int64_t src = d * c->uhat_stride + ((int64_t)j * c->nkz + jz) * c->nkx + kx;
                    const int64_t dst = d * K + (iz * N1 + iy) * Nkx + kx;
                    out[2 * dst] = (double)h[2 * src];
                    out[2 * dst + 1] = (double)h[2 * src + 1];
                }
            }
    return VPBS_OK;
}

}  // namespace

int longrange_get_uhat(vpbs_ctx *c, int32_t *quantity, int64_t *dims3, double *out) {
    if (quantity) *quantity = c->uhat_quantity;
    if (dims3) {
        dims3[0] = c->p.Ns[0] / 2 + 1;
        dims3[1] = c->p.Ns[1];
        dims3[2] = c->p.Ns[2];
    }
    if (!out) return VPBS_OK;
    if (!c->uhat_quantity)
        return fail(c, VPBS_ERR_STATE, "no Fourier field kept: run vpbs_compute with VPBS_KEEP_UHAT (and the long range)");
    return c->p.precision == VPBS_F32 ? get_uhat_impl<float>(c, out) : get_uhat_impl<double>(c, out);
}

int longrange_spectrum(vpbs_ctx *c, int64_t nk, double dk, double *Ek) {
    if (c->uhat_quantity != VPBS_UHAT_VELOCITY && c->uhat_quantity != VPBS_UHAT_STREAMFUNCTION)
        return fail(c, VPBS_ERR_STATE, "no Fourier field kept: run vpbs_compute with VPBS_KEEP_UHAT (and the long range)");
    if (nk < 1 || nk > 8192) return fail(c, VPBS_ERR_INVALID, "spectrum length must be in [1, 8192] (got %lld)", (long long)nk);
    if (!(dk > 0) || !std::isfinite(dk)) return fail(c, VPBS_ERR_INVALID, "wavenumber step must be positive");
    if (!Ek) return fail(c, VPBS_ERR_INVALID, "null output");
    return c->p.precision == VPBS_F32 ? spectrum_impl<float>(c, (int)nk, dk, Ek) : spectrum_impl<double>(c, (int)nk, dk, Ek);
}

int longrange_spectrum_noalloc(vpbs_ctx *c, int64_t nk, double dk, double *Ek) {
    if (c->uhat_quantity != VPBS_UHAT_VELOCITY && c->uhat_quantity != VPBS_UHAT_STREAMFUNCTION)
        return fail(c, VPBS_ERR_STATE, "no Fourier field kept: run vpbs_compute with VPBS_KEEP_UHAT (and the long range)");
    if (nk < 1 || nk > SPEC_NA_MAXBINS)
        return fail(c, VPBS_ERR_INVALID, "spectrum length must be in [1, %d] without allocation (got %lld)",
                    SPEC_NA_MAXBINS, (long long)nk);
    if (!(dk > 0) || !std::isfinite(dk)) return fail(c, VPBS_ERR_INVALID, "wavenumber step must be positive");
    if (!Ek) return fail(c, VPBS_ERR_INVALID, "null output");
    if (!c->spec_scratch) return fail(c, VPBS_ERR_STATE, "scratch address of this GPU not taken");
    return c->p.precision == VPBS_F32 ? spectrum_noalloc_impl<float>(c, (int)nk, dk, Ek)
                                      : spectrum_noalloc_impl<double>(c, (int)nk, dk, Ek);
}

int longrange_band(vpbs_ctx *c, double kmin, double kmax, double *v_out) {
    if (!c->uhat_quantity)
        return fail(c, VPBS_ERR_STATE, "no Fourier field kept: run vpbs_compute with VPBS_KEEP_UHAT (and the long range)");
    if (!(kmin >= 0) || !(kmax >= kmin)) return fail(c, VPBS_ERR_INVALID, "expected 0 <= kmin <= kmax");
    if (!v_out) return fail(c, VPBS_ERR_INVALID, "null output");
    return c->p.precision == VPBS_F32 ? band_impl<float>(c, kmin, kmax, v_out) : band_impl<double>(c, kmin, kmax, v_out);
}

int longrange_filtered_grid(vpbs_ctx *c, double cf, double cc) {
    if (!c->uhat_quantity)
        return fail(c, VPBS_ERR_STATE, "no Fourier field kept: run vpbs_compute with VPBS_KEEP_UHAT (and the long range)");
    const double inf = HUGE_VAL;
    return c->p.precision == VPBS_F32 ? band_grid<float>(c, 0.0, inf, cf, cc) : band_grid<double>(c, 0.0, inf, cf, cc);
}

int longrange_interp_at(vpbs_ctx *c, int64_t n, const double *pos, double *out) {
    if (n <= 0) return VPBS_OK;
    if (c->p.precision == VPBS_F32)
        launch_interp<float>(lr_geom(c), n, pos, (float)c->es_beta, (const float *)c->grid, out, c->s_lr);
    else
        launch_interp<double>(lr_geom(c), n, pos, c->es_beta, (const double *)c->grid, out, c->s_lr);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

void longrange_release(vpbs_ctx *c) {
    spectral_release(c);
    for (auto &pd : c->plan)
        for (auto &p : pd) {
            if (p) rocfft_plan_destroy(p);
            p = nullptr;
        }
    if (c->fft_info) rocfft_execution_info_destroy(c->fft_info);
    c->fft_info = nullptr;
    void *bufs[] = {c->fft_work, c->grid, c->ypen, c->zpen, c->zpen2, c->deconv, c->uhat};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    c->fft_work = c->grid = c->ypen = c->zpen = c->zpen2 = c->uhat = nullptr;
    c->deconv = nullptr;
}

int merge_outputs(vpbs_ctx *c, const double *a, const double *b, double *o, int64_t n, hipStream_t s) {
    add3<<<nblocks(n, 256), 256, 0, s>>>(n, a, b, o);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

}  // namespace vpbs
