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
//   spread_scatter_kernel  one workgroup per 32×8×8 grid tile takes the points of the
//                   3×3×3 neighbouring tiles whose support overlaps it and adds their
//                   kernel-weighted charges into an LDS copy of the tile (each
//                   wavefront owns a z-slab, points in sorted order): every grid value
//                   written exactly once, no global atomics, no zero-fill pass,
//                   deterministic summation order
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

#include <algorithm>
#include <cmath>
#include <vector>

namespace vpbs {

namespace {

constexpr int TX = 32, TY = 8, TZ = 8;  // spread tile (32 cells = one 128-B row in FP32)
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
    int M;
    int64_t cstride;  // real grid values per component
};

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
        t[d] = i / (d == 0 ? TX : (d == 1 ? TY : TZ));
    }
    key[j] = (t[2] * g.ntile[1] + t[1]) * g.ntile[0] + t[0];
}

// Tile-sorted copy of the points with, per point and direction d, the first
// stencil index i0_d = ⌊u_d⌋ - m + 1 (wrapped into [0, G_d)) and the 2m kernel
// weights φ((i0_d + k - u_d)/m), k = 0..2m-1, computed once here instead of
// once per tile in the spreading kernel.
template <typename CT>
__global__ void gather_points(int64_t Nq, const int32_t *__restrict__ perm, const double *__restrict__ pts,
                              const double *__restrict__ chg, LRGeom g, CT beta, int32_t *__restrict__ si0,
                              CT *__restrict__ sw, CT *__restrict__ schg) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Nq) return;
    const int32_t j = perm[s];
    const int M = g.M, W = 2 * M;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double u = pts[d * Nq + j] * g.scale[d];
        const int i0 = (int)floor(u) - M + 1;
        si0[d * Nq + s] = wrapi(i0, g.G[d]);
        for (int k = 0; k < W; k++)
            sw[((int64_t)d * W + k) * Nq + s] = es_kernel<CT>((CT)(((double)(i0 + k) - u) / (double)M), beta);
        schg[d * Nq + s] = (CT)chg[d * Nq + j];
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

// Point-stationary spreading: one workgroup (4 wavefronts) per 32×8×8 tile,
// wavefront w owning the z-slab [2w, 2w+2) of the tile in LDS.
// Per round of up to NB candidates from the 3×3×3 neighbour tiles:
//   A) keep the points whose support overlaps the tile (ballot compaction,
//      candidate order preserved), recording the support start relative to
//      the tile start, wrap1(i0 - t0, G), per direction;
//   B) stage the kept points' 3W weights and charges in LDS;
//   C) each wavefront walks the kept points in order and adds q·wx·wy·wz into
//      the cells of its slab inside the support: lane (a, b) takes support
//      offsets (8i + a, 8j + b) in x and y. A wavefront never shares a cell
//      with another and adds the points in a fixed order, so the sums are
//      deterministic; within one point the lanes hit distinct cells.
// The tile (pair rows interleaved, as in the grid) is then written once.
template <typename CT, int W>
__global__ __launch_bounds__(256) void spread_scatter_kernel(LRGeom g, int64_t Nq, const int32_t *__restrict__ tstart,
                                                             const uint8_t *__restrict__ active,
                                                             const int32_t *__restrict__ si0,
                                                             const CT *__restrict__ sw,
                                                             const CT *__restrict__ schg, CT *__restrict__ grid) {
    constexpr int NB = sizeof(CT) == 4 ? 128 : 64;  // kept points per round
    constexpr int NPAIR = TY / 2;
    constexpr int PITCH = 2 * TX + (sizeof(CT) == 4 ? 16 : 8);  // LDS pitch of a pair row (values)
    constexpr int NPASS = (W + 7) / 8;                           // 8-wide passes per direction
    __shared__ CT acc[3][TZ][NPAIR][PITCH];
    __shared__ CT s_w[NB][3][W];
    __shared__ CT s_q[NB][3];
    __shared__ int4 s_r[NB];  // (rx, ry, rz, sorted point index)
    __shared__ int nb_start[27];
    __shared__ int nb_len[27];
    __shared__ int nb_pre[28];
    __shared__ int s_wave_cnt[4];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int64_t ntl = (int64_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
    const int64_t tile = xcd_tile(blockIdx.x, ntl);
    if (tile >= ntl) return;  // uniform over the workgroup
    const int tix = (int)(tile % g.ntile[0]);
    const int tiy = (int)((tile / g.ntile[0]) % g.ntile[1]);
    const int tiz = (int)(tile / ((int64_t)g.ntile[0] * g.ntile[1]));
    const int t0[3] = {tix * TX, tiy * TY, tiz * TZ};

    CT *accf = &acc[0][0][0][0];
    for (int e = tid; e < 3 * TZ * NPAIR * PITCH; e += 256) accf[e] = (CT)0;

    if (active[tile]) {
        if (tid < 27) {
            const int o[3] = {tid % 3 - 1, (tid / 3) % 3 - 1, tid / 9 - 1};
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
            nb_start[tid] = s0;
            nb_len[tid] = dup ? 0 : tstart[tl + 1] - s0;
        }
        __syncthreads();
        if (tid < 64) {
            int incl = tid < 27 ? nb_len[tid] : 0;
#pragma unroll
            for (int off = 1; off < 32; off <<= 1) {
                int y = __shfl_up(incl, off, 64);
                if (tid >= off) incl += y;
            }
            if (tid < 27) nb_pre[tid + 1] = incl;
            if (tid == 0) nb_pre[0] = 0;
        }
        __syncthreads();
        const int ncand = nb_pre[27];
        const int a8 = lane & 7, b8 = lane >> 3;

        for (int base = 0; base < ncand; base += NB) {
            // A) overlap test and compaction
            bool keep = false;
            int4 rr = {0, 0, 0, 0};
            if (tid < NB && base + tid < ncand) {
                const int cnd = base + tid;
                int lo = 0, hi = 26;  // largest nbi with nb_pre[nbi] <= cnd
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (nb_pre[mid] <= cnd)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                const int sidx = nb_start[lo] + (cnd - nb_pre[lo]);
                int r[3];
                keep = true;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const int i0 = si0[d * Nq + sidx];
                    const int T = d == 0 ? TX : (d == 1 ? TY : TZ);
                    r[d] = wrap1(i0 - t0[d], g.G[d]);                       // support start rel. to tile
                    keep = keep && (r[d] < T || wrap1(t0[d] - i0, g.G[d]) < W);  // overlap (periodic)
                }
                rr = make_int4(r[0], r[1], r[2], sidx);
            }
            const unsigned long long ball = __ballot(keep);
            if (lane == 0) s_wave_cnt[wid] = __popcll(ball);
            __syncthreads();
            if (keep) {
                int pos = __popcll(ball & ((1ull << lane) - 1ull));
                for (int w = 0; w < wid; w++) pos += s_wave_cnt[w];
                s_r[pos] = rr;
            }
            __syncthreads();
            const int ns = s_wave_cnt[0] + s_wave_cnt[1] + s_wave_cnt[2] + s_wave_cnt[3];
            // B) stage weights (component-major source: row d·W + k, point p fastest) and charges
            for (int e = tid; e < ns * 3 * W; e += 256) {
                const int p = e % ns, r = e / ns;
                s_w[p][r / W][r % W] = sw[(int64_t)r * Nq + s_r[p].w];
            }
            for (int e = tid; e < ns * 3; e += 256) {
                const int p = e % ns, d = e / ns;
                s_q[p][d] = schg[(int64_t)d * Nq + s_r[p].w];
            }
            __syncthreads();
            // C) scatter into this wavefront's slab
            const int zs = 2 * wid;
            for (int p = 0; p < ns; p++) {
                const int4 r = s_r[p];
                const int kz0 = wrap1(zs - r.z, g.G[2]), kz1 = wrap1(zs + 1 - r.z, g.G[2]);
                if (kz0 >= W && kz1 >= W) continue;  // uniform: support misses the slab
                const CT q0 = s_q[p][0], q1 = s_q[p][1], q2 = s_q[p][2];
                const CT wz0 = kz0 < W ? s_w[p][2][kz0] : (CT)0;
                const CT wz1 = kz1 < W ? s_w[p][2][kz1] : (CT)0;
#pragma unroll
                for (int pi = 0; pi < NPASS; pi++) {
#pragma unroll
                    for (int pj = 0; pj < NPASS; pj++) {
                        const int kx = 8 * pi + a8, ky = 8 * pj + b8;
                        const int lx = wrap1(r.x + kx, g.G[0]), ly = wrap1(r.y + ky, g.G[1]);
                        if (kx < W && ky < W && lx < TX && ly < TY) {
                            const CT wxy = s_w[p][0][kx] * s_w[p][1][ky];
                            // plain read-modify-write: the slab is this wavefront's alone and
                            // its LDS accesses execute in program order, so the next point's
                            // reads see these writes (LDS float atomics are far slower)
                            CT *cell = &acc[0][zs][ly >> 1][2 * lx + (ly & 1)];
                            constexpr int CS = TZ * NPAIR * PITCH, ZS = NPAIR * PITCH;
                            if (kz0 < W && kz1 < W) {
                                const CT f0 = wxy * wz0, f1 = wxy * wz1;
                                CT a0 = cell[0], a1 = cell[CS], a2 = cell[2 * CS];
                                CT b0 = cell[ZS], b1 = cell[CS + ZS], b2 = cell[2 * CS + ZS];
                                a0 += f0 * q0;
                                a1 += f0 * q1;
                                a2 += f0 * q2;
                                b0 += f1 * q0;
                                b1 += f1 * q1;
                                b2 += f1 * q2;
                                cell[0] = a0;
                                cell[CS] = a1;
                                cell[2 * CS] = a2;
                                cell[ZS] = b0;
                                cell[CS + ZS] = b1;
                                cell[2 * CS + ZS] = b2;
                            } else {
                                const int zo = kz0 < W ? 0 : ZS;
                                const CT f = wxy * (kz0 < W ? wz0 : wz1);
                                CT a0 = cell[zo], a1 = cell[CS + zo], a2 = cell[2 * CS + zo];
                                a0 += f * q0;
                                a1 += f * q1;
                                a2 += f * q2;
                                cell[zo] = a0;
                                cell[CS + zo] = a1;
                                cell[2 * CS + zo] = a2;
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    // write the tile: 3·TZ·NPAIR pair rows of 2·TX contiguous values
    const int v = tid & 63;
    for (int row = tid >> 6; row < 3 * TZ * NPAIR; row += 4) {
        const int d = row / (TZ * NPAIR), zp = row - d * TZ * NPAIR;
        const int z = zp / NPAIR, pp = zp - z * NPAIR;
        const int64_t off = (((int64_t)(t0[2] + z) * (g.G[1] >> 1) + (t0[1] >> 1) + pp) * g.G[0] + t0[0]) * 2 + v;
        grid[d * g.cstride + off] = acc[d][z][pp][v];
    }
}

template <typename CT>
void launch_spread(const LRGeom &g, int64_t nb, int64_t Nq, const int32_t *tstart, const uint8_t *active,
                   const int32_t *si0, const CT *sw, const CT *schg, CT *grid, hipStream_t s) {
    const unsigned n = (unsigned)nb;
    switch (g.M) {
        case 2: spread_scatter_kernel<CT, 4><<<n, 256, 0, s>>>(g, Nq, tstart, active, si0, sw, schg, grid); break;
        case 3: spread_scatter_kernel<CT, 6><<<n, 256, 0, s>>>(g, Nq, tstart, active, si0, sw, schg, grid); break;
        case 4: spread_scatter_kernel<CT, 8><<<n, 256, 0, s>>>(g, Nq, tstart, active, si0, sw, schg, grid); break;
        case 5: spread_scatter_kernel<CT, 10><<<n, 256, 0, s>>>(g, Nq, tstart, active, si0, sw, schg, grid); break;
        case 6: spread_scatter_kernel<CT, 12><<<n, 256, 0, s>>>(g, Nq, tstart, active, si0, sw, schg, grid); break;
        case 7: spread_scatter_kernel<CT, 14><<<n, 256, 0, s>>>(g, Nq, tstart, active, si0, sw, schg, grid); break;
        default: spread_scatter_kernel<CT, 16><<<n, 256, 0, s>>>(g, Nq, tstart, active, si0, sw, schg, grid); break;
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
    if (!keep || k2 == 0) {
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
// Interpolation (type 2).

// Type-2 interpolation with a W³ stencil (W = 2m): one wavefront per node.
// Lane (a, b) = (lane % W, lane / W) owns stencil column x = i0x + a and rows
// y = i0y + b + R·pass (R = ⌊64/W⌋ rows per pass) and walks the W z-planes, so a
// load instruction touches R rows of W consecutive values; the partial sums
// are combined by a butterfly reduction (pairwise summation). Every lane
// evaluates the W z-weights itself: broadcasting FP64 weights with a wave
// shuffle was measured to lose precision (~1e-7 relative against the oracle),
// and the redundant evaluation is cheap next to the grid loads.
template <typename CT, int W>
__global__ __launch_bounds__(256) void interp_kernel_wave(LRGeom g, int64_t Np, const double *__restrict__ nodes,
                                                          CT beta, const CT *__restrict__ grid,
                                                          double *__restrict__ out) {
    constexpr int M = W / 2;
    constexpr int R = 64 / W;                  // y rows per pass
    constexpr int PASSES = (W + R - 1) / R;
    const int lane = threadIdx.x & 63;
    const int a = lane % W, b = lane / W;
    const bool lane_on = b < R;                // lanes beyond R·W idle
    // Blocks are dealt to the 8 XCDs round-robin; remap so that each XCD works
    // on a contiguous range of nodes (consecutive nodes share stencil rows,
    // which then stay in that XCD's L2).
    const int64_t per = ((int64_t)gridDim.x + 7) / 8;
    const int64_t blk = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    {
        const int64_t i = blk * 4 + (threadIdx.x >> 6);
        if (i >= Np) return;
        const double ux = nodes[i] * g.scale[0];
        const double uy = nodes[Np + i] * g.scale[1];
        const double uz = nodes[2 * Np + i] * g.scale[2];
        const int i0x = (int)floor(ux) - M + 1, i0y = (int)floor(uy) - M + 1, i0z = (int)floor(uz) - M + 1;
        const CT wx = es_kernel<CT>((CT)(((double)(i0x + a) - ux) / (double)M), beta);
        CT wz[W];
        int64_t gzo[W];
        const int64_t plane = (int64_t)g.G[0] * g.G[1];  // values per z-plane
#pragma unroll
        for (int c = 0; c < W; c++) {
            wz[c] = es_kernel<CT>((CT)(((double)(i0z + c) - uz) / (double)M), beta);
            gzo[c] = (int64_t)wrapi(i0z + c, g.G[2]) * plane;
        }
        const int64_t gx = 2 * (int64_t)wrapi(i0x + a, g.G[0]);
        CT acc0 = 0, acc1 = 0, acc2 = 0;
#pragma unroll
        for (int pass = 0; pass < PASSES; pass++) {
            const int bb = b + R * pass;
            if (lane_on && bb < W) {
                const CT wy = es_kernel<CT>((CT)(((double)(i0y + bb) - uy) / (double)M), beta);
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
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            acc0 += __shfl_xor(acc0, o, 64);
            acc1 += __shfl_xor(acc1, o, 64);
            acc2 += __shfl_xor(acc2, o, 64);
        }
        if (lane == 0) {
            out[i] = (double)acc0;
            out[Np + i] = (double)acc1;
            out[2 * Np + i] = (double)acc2;
        }
    }
}

template <typename CT>
void launch_interp(const LRGeom &g, int64_t Np, const double *nodes, CT beta, const CT *grid, double *out,
                   hipStream_t s) {
    const unsigned nb = (unsigned)(8 * (((Np + 3) / 4 + 7) / 8));  // one wavefront per node, 4 per block
    switch (g.M) {
        case 2: interp_kernel_wave<CT, 4><<<nb, 256, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        case 3: interp_kernel_wave<CT, 6><<<nb, 256, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        case 4: interp_kernel_wave<CT, 8><<<nb, 256, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        case 5: interp_kernel_wave<CT, 10><<<nb, 256, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        case 6: interp_kernel_wave<CT, 12><<<nb, 256, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        case 7: interp_kernel_wave<CT, 14><<<nb, 256, 0, s>>>(g, Np, nodes, beta, grid, out); break;
        default: interp_kernel_wave<CT, 16><<<nb, 256, 0, s>>>(g, Np, nodes, beta, grid, out); break;
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
int launch_lr(vpbs_ctx *c, uint32_t flags) {
    hipStream_t s = c->s_lr;
    const int64_t Np = c->Np, Nq = c->Nq;
    const bool vel = flags & VPBS_VELOCITY, psi = flags & VPBS_STREAMFUNCTION;
    LRGeom g;
    for (int d = 0; d < 3; d++) {
        g.G[d] = c->G[d];
        g.scale[d] = (double)c->G[d] / c->p.Ls[d];
        g.ntile[d] = c->ntile[d];
    }
    g.M = c->p.nufft_m;
    g.cstride = c->comp_stride;
    const CT beta = (CT)c->es_beta;

    stage_begin(c, ST_LRSORT, s);
    lr_tile_keys<<<nblocks(Nq, 256), 256, 0, s>>>(Nq, c->points, g, c->lr_key);
    VPBS_TRY(counting_sort(c, c->sort_lr, c->lr_key, Nq, c->ntiles, c->lr_start, c->lr_perm, s));
    gather_points<CT><<<nblocks(Nq, 256), 256, 0, s>>>(Nq, c->lr_perm, c->points, c->charges, g, beta, c->lr_i0,
                                                        (CT *)c->lr_w, (CT *)c->lr_chg);
    stage_end(c, ST_LRSORT, s);

    CT *grid = (CT *)c->grid;
    stage_begin(c, ST_SPREAD, s);
    tile_active<<<nblocks(c->ntiles, 256), 256, 0, s>>>(g, c->ntiles, c->lr_start, c->lr_active);
    const int64_t nb_spread = ((c->ntiles + 7) / 8) * 8;  // xcd_tile() maps 8·⌈n/8⌉ blocks onto the tiles
    launch_spread<CT>(g, nb_spread, Nq, c->lr_start, c->lr_active, c->lr_i0, (const CT *)c->lr_w,
                      (const CT *)c->lr_chg, grid, s);
    stage_end(c, ST_SPREAD, s);
    VPBS_HIP(c, hipGetLastError());

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
    // z-pencils -> ypen -> pair rows -> real grid in place
    auto inverse = [&](Cx<CT> *zp) -> int {
        VPBS_TRY(fft(2, 1, zp));
        zpencils_to_ypencils<CT><<<dim3(G2 / 32, G1 / 32, 3 * nkx), 256, 0, s>>>(zp, ypen, G1, G2, nky, K1);
        VPBS_TRY(fft(1, 1, ypen));
        ypencils_to_pair_rows<CT><<<dim3(G0 / 32, G1 / 32, 3 * G2), 256, 0, s>>>(ypen, rows, G0, G1, G2, nkx);
        VPBS_TRY(fft(0, 1, rows));
        VPBS_HIP(c, hipGetLastError());
        return VPBS_OK;
    };

    stage_begin(c, ST_FFTF, s);
    Cx<CT> *zpen = (Cx<CT> *)c->zpen;
    VPBS_TRY(fft(0, 0, rows));
    pair_split_to_ypencils<CT><<<dim3((nkx + 31) / 32, G1 / 32, 3 * G2), 256, 0, s>>>(rows, ypen, G0, G1, G2, nkx);
    VPBS_TRY(fft(1, 0, ypen));
    ypencils_to_zpencils<CT><<<dim3(G2 / 32, (nky + 31) / 32, 3 * nkx), 256, 0, s>>>(ypen, zpen, G1, G2, nky, K1);
    VPBS_TRY(fft(2, 0, zpen));
    stage_end(c, ST_FFTF, s);
    VPBS_HIP(c, hipGetLastError());

    FourierArgs<CT> fa;
    for (int d = 0; d < 3; d++) {
        fa.G[d] = c->G[d];
        fa.kmax[d] = c->kmax_mode[d];
        fa.kscale[d] = 2 * M_PI / c->p.Ls[d];
    }
    fa.nkx = nkx;
    fa.nky = nky;
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
    fa.do_vel = vel;
    fa.do_psi = psi;
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

int longrange_setup(vpbs_ctx *c) {
    const int m = c->p.nufft_m;
    for (int d = 0; d < 3; d++) {
        c->G[d] = oversampled_size(c->p.Ns[d], c->p.nufft_sigma);
        if (c->G[d] < 4 * m) c->G[d] = ((4 * m + 31) / 32) * 32;
        c->ntile[d] = c->G[d] / (d == 0 ? TX : (d == 1 ? TY : TZ));
    }
    c->ntiles = (int64_t)c->ntile[0] * c->ntile[1] * c->ntile[2];
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
    if (c->p.precision == VPBS_F32) return launch_lr<float>(c, flags);
    return launch_lr<double>(c, flags);
}

void longrange_release(vpbs_ctx *c) {
    for (auto &pd : c->plan)
        for (auto &p : pd) {
            if (p) rocfft_plan_destroy(p);
            p = nullptr;
        }
    if (c->fft_info) rocfft_execution_info_destroy(c->fft_info);
    c->fft_info = nullptr;
    void *bufs[] = {c->fft_work, c->grid, c->ypen, c->zpen, c->zpen2, c->deconv};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    c->fft_work = c->grid = c->ypen = c->zpen = c->zpen2 = nullptr;
    c->deconv = nullptr;
}

int merge_outputs(vpbs_ctx *c, const double *a, const double *b, double *o, int64_t n, hipStream_t s) {
    add3<<<nblocks(n, 256), 256, 0, s>>>(n, a, b, o);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

}  // namespace vpbs
