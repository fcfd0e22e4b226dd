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
//   spread_kernel   one workgroup per 16×8×8 grid tile GATHERS the kernel-weighted
//                   charges of the points in the 3×3×3 neighbouring tiles: weights
//                   staged in LDS, every grid value written exactly once, no atomics,
//                   no zero-fill pass, deterministic summation order
//   fourier_kernel  deconvolution (twice: type 1 and type 2), Γ/V, Green function,
//                   curl, Ewald Gaussian, Hermitian clean-up, zeroing of dropped
//                   modes — one pass over the half spectrum, in place
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
    int64_t Gxp;
    int64_t cstride;
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

// One workgroup (256 threads) per 32×8×8 tile. Wave w owns the x-octet
// x ∈ [8w, 8w+8) of the tile and lane l the row (y, z) = (l % 8, l / 8); each
// thread accumulates its 8 cells × 3 components in registers.
// For each batch of up to SPB candidate points from the 3×3×3 neighbour tiles:
//   A) keep the points whose kernel support overlaps the tile (deterministic
//      ballot compaction, candidate order preserved),
//   B) expand their precomputed weights to per-tile rows in LDS (zero outside
//      the support) and record which x-octets the support touches,
//   C) accumulate Σ_p q_p·wx·wy·wz; a wave skips the points whose support
//      misses its octet (a wave-uniform branch).
// The tile is then transposed through LDS and written once, in full rows.
constexpr int SPREAD_TPB = 256;
constexpr int SPB = 128;

template <typename CT>
struct SpreadSmem {
    CT wx[SPB][TX];
    CT wy[SPB][TY];
    CT wz[SPB][TZ];
};

template <typename CT>
__global__ __launch_bounds__(SPREAD_TPB) void spread_kernel(LRGeom g, int64_t Nq, const int32_t *__restrict__ tstart,
                                                            const uint8_t *__restrict__ active,
                                                            const int32_t *__restrict__ si0,
                                                            const CT *__restrict__ sw,
                                                            const CT *__restrict__ schg, CT *__restrict__ grid) {
    static_assert(sizeof(SpreadSmem<CT>) >= 3 * TX * TY * TZ * sizeof(CT), "epilogue tile must fit");
    __shared__ int nb_start[27];
    __shared__ int nb_len[27];
    __shared__ int nb_pre[28];
    __shared__ SpreadSmem<CT> sm;
    __shared__ Q4<CT> s_q[SPB];
    __shared__ int s_pt[SPB];  // staged point index (sorted order)
    __shared__ int s_mask[SPB];  // x-octets of the tile touched by the point's support
    __shared__ int s_wave_cnt[SPB / 64];

    const int tid = threadIdx.x;
    const int64_t ntl = (int64_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
    const int64_t tile = xcd_tile(blockIdx.x, ntl);
    if (tile >= ntl) return;
    const int tix = (int)(tile % g.ntile[0]);
    const int tiy = (int)((tile / g.ntile[0]) % g.ntile[1]);
    const int tiz = (int)(tile / ((int64_t)g.ntile[0] * g.ntile[1]));
    const int t0[3] = {tix * TX, tiy * TY, tiz * TZ};
    const int lane = tid & 63, wid = tid >> 6;
    const int ly = lane & 7, lz = lane >> 3;  // this thread's row

    CT acc[8][3];
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
        for (int d = 0; d < 3; d++) acc[a][d] = (CT)0;

    if (active[tile]) {
        const int W = 2 * g.M;
        // Neighbour tiles t + (-1..1)³: lane k < 27 loads the point range of one
        // of them; a direction with fewer than 3 tiles would visit a tile twice,
        // so the repeated offsets get an empty range.
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

        for (int base = 0; base < ncand; base += SPB) {
            // A) candidate -> overlap test (support [i0, i0+W) ∩ tile, periodic) -> compaction
            bool keep = false;
            int sidx = 0, mask = 0;
            if (tid < SPB && base + tid < ncand) {
                const int c = base + tid;
                int lo = 0, hi = 26;  // largest nbi with nb_pre[nbi] <= c
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (nb_pre[mid] <= c)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                sidx = nb_start[lo] + (c - nb_pre[lo]);
                keep = true;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const int i0 = si0[d * Nq + sidx];
                    const int T = d == 0 ? TX : (d == 1 ? TY : TZ);
                    const int dd = wrap1(t0[d] - i0, g.G[d]);  // tile start relative to support start
                    const int ee = wrap1(i0 - t0[d], g.G[d]);  // support start relative to tile start
                    keep = keep && (dd < W || ee < T);
                    if (d == 0) {
#pragma unroll
                        for (int o = 0; o < TX / 8; o++) {
                            const int d8 = wrap1(t0[0] + 8 * o - i0, g.G[0]);
                            const int e8 = wrap1(i0 - t0[0] - 8 * o + g.G[0], g.G[0]);
                            mask |= (d8 < W || e8 < 8) ? (1 << o) : 0;
                        }
                    }
                }
            }
            const unsigned long long ball = __ballot(keep);
            if (tid < SPB && lane == 0) s_wave_cnt[wid] = __popcll(ball);
            __syncthreads();
            if (keep) {
                int pos = __popcll(ball & ((1ull << lane) - 1ull));
                for (int w = 0; w < wid; w++) pos += s_wave_cnt[w];
                s_pt[pos] = sidx;
                s_mask[pos] = mask;
                s_q[pos] = Q4<CT>{schg[sidx], schg[Nq + sidx], schg[2 * Nq + sidx], (CT)0};
            }
            __syncthreads();
            int ns = 0;
#pragma unroll
            for (int w = 0; w < SPB / 64; w++) ns += s_wave_cnt[w];
            // B) expand the weights of the staged points over the tile rows
            for (int e = tid; e < ns * (TX + TY + TZ); e += SPREAD_TPB) {
                const int p = e / (TX + TY + TZ);
                const int r = e - p * (TX + TY + TZ);
                const int d = r < TX ? 0 : (r < TX + TY ? 1 : 2);
                const int cl = d == 0 ? r : (d == 1 ? r - TX : r - TX - TY);
                const int sp = s_pt[p];
                const int k = wrap1(t0[d] + cl - si0[d * Nq + sp], g.G[d]);
                const CT w = k < W ? sw[((int64_t)d * W + k) * Nq + sp] : (CT)0;
                if (d == 0)
                    sm.wx[p][cl] = w;
                else if (d == 1)
                    sm.wy[p][cl] = w;
                else
                    sm.wz[p][cl] = w;
            }
            __syncthreads();
            // C) accumulate
            const int bit = 1 << wid;
            for (int p = 0; p < ns; p++) {
                if (!(s_mask[p] & bit)) continue;  // wave-uniform
                const CT wyz = sm.wy[p][ly] * sm.wz[p][lz];
                const Q4<CT> q = s_q[p];
                const CT a0 = wyz * q.x, a1 = wyz * q.y, a2 = wyz * q.z;
#pragma unroll
                for (int a = 0; a < 8; a++) {
                    const CT wx = sm.wx[p][8 * wid + a];
                    acc[a][0] += wx * a0;
                    acc[a][1] += wx * a1;
                    acc[a][2] += wx * a2;
                }
            }
            __syncthreads();
        }
    }
    // Transpose through LDS and store whole rows: tile[d][z][y][x].
    CT *tilebuf = reinterpret_cast<CT *>(&sm);
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
        for (int d = 0; d < 3; d++) tilebuf[((d * TZ + lz) * TY + ly) * TX + 8 * wid + a] = acc[a][d];
    __syncthreads();
    const int x = tid & 31;
    for (int r = tid >> 5; r < 3 * TY * TZ; r += SPREAD_TPB / 32) {
        const int d = r / (TY * TZ), yz = r - d * TY * TZ;
        const int z = yz / TY, y = yz - z * TY;
        const int64_t off = ((int64_t)(t0[2] + z) * g.G[1] + (t0[1] + y)) * g.Gxp + t0[0] + x;
        grid[d * g.cstride + off] = tilebuf[r * TX + x];
    }
}

// ----------------------------------------------------------------------------
// Fused Fourier-space kernel (in place on the half spectrum).

template <typename CT>
struct FourierArgs {
    int G[3];
    int64_t Hx;       // G0/2 + 1
    int64_t cstride;  // complex elements per component
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
    const int64_t total = a.Hx * a.G[1] * a.G[2];
    if (idx >= total) return;
    const int ix = (int)(idx % a.Hx);
    const int iy = (int)((idx / a.Hx) % a.G[1]);
    const int iz = (int)(idx / (a.Hx * a.G[1]));
    const int kx = ix;
    const int ky = iy <= a.G[1] / 2 ? iy : iy - a.G[1];
    const int kz = iz <= a.G[2] / 2 ? iz : iz - a.G[2];
    bool keep = kx <= a.kmax[0] && ky >= -a.kmax[1] && ky <= a.kmax[1] && kz >= -a.kmax[2] && kz <= a.kmax[2];
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
#pragma unroll
        for (int c = 0; c < W; c++) {
            wz[c] = es_kernel<CT>((CT)(((double)(i0z + c) - uz) / (double)M), beta);
            gzo[c] = (int64_t)wrapi(i0z + c, g.G[2]) * g.G[1];
        }
        const int64_t gx = wrapi(i0x + a, g.G[0]);
        CT acc0 = 0, acc1 = 0, acc2 = 0;
#pragma unroll
        for (int pass = 0; pass < PASSES; pass++) {
            const int bb = b + R * pass;
            if (lane_on && bb < W) {
                const CT wy = es_kernel<CT>((CT)(((double)(i0y + bb) - uy) / (double)M), beta);
                const int64_t gy = wrapi(i0y + bb, g.G[1]);
                const CT wxy = wx * wy;
#pragma unroll
                for (int c = 0; c < W; c++) {
                    const int64_t off = (gzo[c] + gy) * g.Gxp + gx;
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
    g.Gxp = c->Gxp;
    g.cstride = c->comp_stride;
    const CT beta = (CT)c->es_beta;

    stage_begin(c, ST_LRSORT, s);
    lr_tile_keys<<<nblocks(Nq, 256), 256, 0, s>>>(Nq, c->points, g, c->lr_key);
    VPBS_TRY(counting_sort(c, c->sort_lr, c->lr_key, Nq, c->ntiles, c->lr_start, c->lr_perm, s));
    gather_points<CT><<<nblocks(Nq, 256), 256, 0, s>>>(Nq, c->lr_perm, c->points, c->charges, g, beta, c->lr_i0,
                                                        (CT *)c->lr_w, (CT *)c->lr_chg);
    stage_end(c, ST_LRSORT, s);

    CT *grid = (CT *)c->grid;
    CT *spec = (CT *)c->spec;
    CT *spec2 = (CT *)c->spec2;
    stage_begin(c, ST_SPREAD, s);
    tile_active<<<nblocks(c->ntiles, 256), 256, 0, s>>>(g, c->ntiles, c->lr_start, c->lr_active);
    const int64_t nb_spread = ((c->ntiles + 7) / 8) * 8;  // xcd_tile() maps 8·⌈n/8⌉ blocks onto the tiles
    spread_kernel<CT><<<(unsigned)nb_spread, SPREAD_TPB, 0, s>>>(g, Nq, c->lr_start, c->lr_active, c->lr_i0,
                                                                  (const CT *)c->lr_w, (const CT *)c->lr_chg, grid);
    stage_end(c, ST_SPREAD, s);
    VPBS_HIP(c, hipGetLastError());

    stage_begin(c, ST_FFTF, s);
    void *fin[1] = {grid}, *fout[1] = {spec};
    if (rocfft_execute(c->plan_fwd, fin, fout, c->info_fwd) != rocfft_status_success)
        return fail(c, VPBS_ERR_FFT, "rocfft_execute (forward) failed");
    stage_end(c, ST_FFTF, s);

    FourierArgs<CT> fa;
    for (int d = 0; d < 3; d++) {
        fa.G[d] = c->G[d];
        fa.kmax[d] = c->kmax_mode[d];
        fa.kscale[d] = 2 * M_PI / c->p.Ls[d];
    }
    fa.Hx = c->G[0] / 2 + 1;
    fa.cstride = c->spec_stride;
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
    const int64_t nspec = fa.Hx * c->G[1] * c->G[2];
    stage_begin(c, ST_FOURIER, s);
    fourier_kernel<CT><<<nblocks(nspec, 256), 256, 0, s>>>(fa, spec, psi ? spec2 : spec);
    stage_end(c, ST_FOURIER, s);
    VPBS_HIP(c, hipGetLastError());

    // velocity: spectrum -> real grid -> nodes; then the streamfunction reuses the grid
    stage_begin(c, ST_FFTB, s);
    if (vel) {
        void *bin[1] = {spec}, *bout[1] = {grid};
        if (rocfft_execute(c->plan_bwd, bin, bout, c->info_bwd) != rocfft_status_success)
            return fail(c, VPBS_ERR_FFT, "rocfft_execute (backward) failed");
    }
    stage_end(c, ST_FFTB, s);
    stage_begin(c, ST_INTERP, s);
    if (vel) launch_interp<CT>(g, Np, c->nodes, beta, grid, c->v_lr, s);
    if (psi) {
        void *bin[1] = {spec2}, *bout[1] = {grid};
        if (rocfft_execute(c->plan_bwd, bin, bout, c->info_bwd) != rocfft_status_success)
            return fail(c, VPBS_ERR_FFT, "rocfft_execute (backward, psi) failed");
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
    c->Gxp = c->G[0];
    c->comp_stride = (int64_t)c->G[0] * c->G[1] * c->G[2];
    c->spec_stride = (int64_t)(c->G[0] / 2 + 1) * c->G[1] * c->G[2];
    c->es_beta = es_beta(m, c->p.nufft_sigma);
    // retained modes (longrange.jl:556-583)
    c->kmax_mode[0] = c->p.Ns[0] / 2 - 1;
    for (int d = 1; d < 3; d++)
        c->kmax_mode[d] = (c->p.Ns[d] % 2 == 0) ? c->p.Ns[d] / 2 - 1 : (c->p.Ns[d] - 1) / 2;

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

    // grid buffers: real grid and half spectrum (out-of-place transforms keep the
    // real rows unpadded, i.e. aligned to 128 B for spreading and interpolation)
    const size_t es = c->p.precision == VPBS_F32 ? sizeof(float) : sizeof(double);
    VPBS_HIP(c, hipMalloc(&c->grid, es * 3 * (size_t)c->comp_stride));
    VPBS_HIP(c, hipMalloc(&c->spec, 2 * es * 3 * (size_t)c->spec_stride));

    // rocFFT plans: 3 transforms (components) of G0×G1×G2, x fastest, out of place
    rocfft_precision prec = c->p.precision == VPBS_F32 ? rocfft_precision_single : rocfft_precision_double;
    size_t len[3] = {(size_t)c->G[0], (size_t)c->G[1], (size_t)c->G[2]};
    const size_t Hx = c->G[0] / 2 + 1;
    size_t rstr[3] = {1, (size_t)c->Gxp, (size_t)c->Gxp * c->G[1]};
    size_t cstr[3] = {1, Hx, Hx * c->G[1]};
    size_t rdist = (size_t)c->comp_stride, cdist = (size_t)c->spec_stride;
    rocfft_plan_description df = nullptr, db = nullptr;
    if (!rocfft_ok(rocfft_plan_description_create(&df)) || !rocfft_ok(rocfft_plan_description_create(&db)))
        return fail(c, VPBS_ERR_FFT, "rocfft_plan_description_create failed");
    if (!rocfft_ok(rocfft_plan_description_set_data_layout(df, rocfft_array_type_real,
                                                           rocfft_array_type_hermitian_interleaved, nullptr,
                                                           nullptr, 3, rstr, rdist, 3, cstr, cdist)) ||
        !rocfft_ok(rocfft_plan_description_set_data_layout(db, rocfft_array_type_hermitian_interleaved,
                                                           rocfft_array_type_real, nullptr, nullptr, 3, cstr,
                                                           cdist, 3, rstr, rdist)))
        return fail(c, VPBS_ERR_FFT, "rocfft_plan_description_set_data_layout failed");
    if (!rocfft_ok(rocfft_plan_create(&c->plan_fwd, rocfft_placement_notinplace,
                                      rocfft_transform_type_real_forward, prec, 3, len, 3, df)) ||
        !rocfft_ok(rocfft_plan_create(&c->plan_bwd, rocfft_placement_notinplace,
                                      rocfft_transform_type_real_inverse, prec, 3, len, 3, db)))
        return fail(c, VPBS_ERR_FFT, "rocfft_plan_create failed (grid %d x %d x %d)", c->G[0], c->G[1], c->G[2]);
    rocfft_plan_description_destroy(df);
    rocfft_plan_description_destroy(db);
    size_t wf = 0, wb = 0;
    rocfft_plan_get_work_buffer_size(c->plan_fwd, &wf);
    rocfft_plan_get_work_buffer_size(c->plan_bwd, &wb);
    c->fft_work_size = wf > wb ? wf : wb;
    if (c->fft_work_size) VPBS_HIP(c, hipMalloc(&c->fft_work, c->fft_work_size));
    if (!rocfft_ok(rocfft_execution_info_create(&c->info_fwd)) ||
        !rocfft_ok(rocfft_execution_info_create(&c->info_bwd)))
        return fail(c, VPBS_ERR_FFT, "rocfft_execution_info_create failed");
    rocfft_execution_info_set_stream(c->info_fwd, c->s_lr);
    rocfft_execution_info_set_stream(c->info_bwd, c->s_lr);
    if (c->fft_work_size) {
        rocfft_execution_info_set_work_buffer(c->info_fwd, c->fft_work, c->fft_work_size);
        rocfft_execution_info_set_work_buffer(c->info_bwd, c->fft_work, c->fft_work_size);
    }
    c->plans_ready = true;
    return VPBS_OK;
}

int longrange_compute(vpbs_ctx *c, uint32_t flags) {
    const bool psi = flags & VPBS_STREAMFUNCTION;
    const size_t es = c->p.precision == VPBS_F32 ? sizeof(float) : sizeof(double);
    if (psi && !c->spec2) VPBS_HIP(c, hipMalloc(&c->spec2, 2 * es * 3 * (size_t)c->spec_stride));
    if (c->p.precision == VPBS_F32) return launch_lr<float>(c, flags);
    return launch_lr<double>(c, flags);
}

void longrange_release(vpbs_ctx *c) {
    if (c->plan_fwd) rocfft_plan_destroy(c->plan_fwd);
    if (c->plan_bwd) rocfft_plan_destroy(c->plan_bwd);
    if (c->info_fwd) rocfft_execution_info_destroy(c->info_fwd);
    if (c->info_bwd) rocfft_execution_info_destroy(c->info_bwd);
    if (c->fft_work) (void)hipFree(c->fft_work);
    if (c->grid) (void)hipFree(c->grid);
    if (c->spec) (void)hipFree(c->spec);
    if (c->spec2) (void)hipFree(c->spec2);
    if (c->deconv) (void)hipFree(c->deconv);
    c->plan_fwd = c->plan_bwd = nullptr;
    c->info_fwd = c->info_bwd = nullptr;
    c->fft_work = c->grid = c->spec = c->spec2 = nullptr;
    c->deconv = nullptr;
}

int merge_outputs(vpbs_ctx *c, const double *a, const double *b, double *o, int64_t n, hipStream_t s) {
    add3<<<nblocks(n, 256), 256, 0, s>>>(n, a, b, o);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

}  // namespace vpbs
