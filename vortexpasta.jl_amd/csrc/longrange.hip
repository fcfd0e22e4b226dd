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
//   interp_kernel   8×8×8 ES-weighted gather per node
// Everything is templated on the grid type (double for the FP64 path, float
// for the FP32-bulk path).
#include "vpbs_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace vpbs {

namespace {

constexpr int TX = 16, TY = 8, TZ = 8;  // spread tile
constexpr int SPREAD_THREADS = 256;
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

template <typename CT>
__global__ void gather_points(int64_t Nq, const int32_t *__restrict__ perm, const double *__restrict__ pts,
                              const double *__restrict__ chg, double *__restrict__ spos, CT *__restrict__ schg) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Nq) return;
    int32_t j = perm[s];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        spos[d * Nq + s] = pts[d * Nq + j];
        schg[d * Nq + s] = (CT)chg[d * Nq + j];
    }
}

// ----------------------------------------------------------------------------
// Spreading (type 1), gather form.

template <typename CT>
__global__ __launch_bounds__(SPREAD_THREADS) void spread_kernel(LRGeom g, int64_t Nq, const int32_t *__restrict__ tstart,
                                                               const double *__restrict__ spos,
                                                               const CT *__restrict__ schg, CT beta,
                                                               CT *__restrict__ grid) {
    constexpr int PB = sizeof(CT) == 4 ? 256 : 128;  // staged points per batch
    __shared__ int nb_start[27];
    __shared__ int nb_len[27];
    __shared__ int nb_pre[28];
    __shared__ CT s_wx[PB][TX + 1];
    __shared__ CT s_wy[PB][TY + 1];
    __shared__ CT s_wz[PB][TZ + 1];
    __shared__ CT s_q[PB][3];
    __shared__ double s_u[PB][3];
    __shared__ int s_i0[PB][3];
    __shared__ int s_wave_cnt[PB / 64];
    __shared__ int s_nstaged;

    const int tid = threadIdx.x;
    const int64_t tile = blockIdx.x;
    const int tix = (int)(tile % g.ntile[0]);
    const int tiy = (int)((tile / g.ntile[0]) % g.ntile[1]);
    const int tiz = (int)(tile / ((int64_t)g.ntile[0] * g.ntile[1]));
    const int t0[3] = {tix * TX, tiy * TY, tiz * TZ};
    const int M = g.M;
    const int W = 2 * M;

    // Neighbour tiles t + (-1..1)³: lane k < 27 loads the point range of one of
    // them; a direction with fewer than 3 tiles would visit a tile twice, so the
    // repeated offsets get an empty range.
    if (tid < 27) {
        const int o[3] = {tid % 3 - 1, (tid / 3) % 3 - 1, tid / 9 - 1};
        const int tt[3] = {tix, tiy, tiz};
        int tn[3];
        bool dup = false;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            tn[d] = wrapi(tt[d] + o[d], g.ntile[d]);
            for (int op = -1; op < o[d]; op++) dup |= (wrapi(tt[d] + op, g.ntile[d]) == tn[d]);
        }
        const int64_t tl = ((int64_t)tn[2] * g.ntile[1] + tn[1]) * g.ntile[0] + tn[0];
        const int s0 = tstart[tl];
        nb_start[tid] = s0;
        nb_len[tid] = dup ? 0 : tstart[tl + 1] - s0;
    }
    __syncthreads();
    if (tid < 64) {
        int len = tid < 27 ? nb_len[tid] : 0;
        int incl = len;
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

    // cells of this thread: x = tid % 16, y = (tid / 16) % 8, z = tid / 128 + 2m
    const int cxl = tid & 15, cyl = (tid >> 4) & 7, czl0 = tid >> 7;
    CT acc[4][3];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int d = 0; d < 3; d++) acc[m][d] = (CT)0;

    const int lane = tid & 63, wid = tid >> 6;
    for (int base = 0; base < ncand; base += PB) {
        // Phase A: candidates -> overlap test -> deterministic compaction
        bool keep = false;
        double u[3];
        int i0[3];
        int64_t sidx = 0;
        if (tid < PB && base + tid < ncand) {
            const int c = base + tid;
            int lo = 0, hi = 26;  // largest nbi with nb_pre[nbi] <= c
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (nb_pre[mid] <= c) lo = mid; else hi = mid - 1;
            }
            sidx = nb_start[lo] + (c - nb_pre[lo]);
            keep = true;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                u[d] = spos[d * Nq + sidx] * g.scale[d];
                i0[d] = (int)floor(u[d]) - M + 1;
                const int T = d == 0 ? TX : (d == 1 ? TY : TZ);
                int dd = wrapi(t0[d] - i0[d], g.G[d]);  // tile start relative to support start
                int ee = wrapi(i0[d] - t0[d], g.G[d]);  // support start relative to tile start
                keep = keep && (dd < W || ee < T);
            }
        }
        unsigned long long ball = __ballot(keep);
        if (tid < PB && lane == 0) s_wave_cnt[wid] = __popcll(ball);
        __syncthreads();
        if (keep) {
            int pos = __popcll(ball & ((1ull << lane) - 1ull));
            for (int w = 0; w < wid; w++) pos += s_wave_cnt[w];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                s_u[pos][d] = u[d];
                s_i0[pos][d] = i0[d];
                s_q[pos][d] = schg[d * Nq + sidx];
            }
        }
        if (tid == 0) {
            int n = 0;
            for (int w = 0; w < PB / 64; w++) n += s_wave_cnt[w];
            s_nstaged = n;
        }
        __syncthreads();
        const int ns = s_nstaged;
        // Phase B: per-tile weights (zero outside the kernel support)
        for (int e = tid; e < ns * (TX + TY + TZ); e += SPREAD_THREADS) {
            int p = e / (TX + TY + TZ);
            int r = e - p * (TX + TY + TZ);
            int d = r < TX ? 0 : (r < TX + TY ? 1 : 2);
            int cl = d == 0 ? r : (d == 1 ? r - TX : r - TX - TY);
            int k = wrapi(t0[d] + cl - s_i0[p][d], g.G[d]);
            CT w = (CT)0;
            if (k < W) {
                double zz = ((double)(s_i0[p][d] + k) - s_u[p][d]) / (double)M;
                w = es_kernel<CT>((CT)zz, beta);
            }
            if (d == 0)
                s_wx[p][cl] = w;
            else if (d == 1)
                s_wy[p][cl] = w;
            else
                s_wz[p][cl] = w;
        }
        __syncthreads();
        // Phase C: accumulate
        for (int p = 0; p < ns; p++) {
            const CT wxy = s_wx[p][cxl] * s_wy[p][cyl];
            const CT q0 = s_q[p][0], q1 = s_q[p][1], q2 = s_q[p][2];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const CT w = wxy * s_wz[p][czl0 + 2 * m];
                acc[m][0] += w * q0;
                acc[m][1] += w * q1;
                acc[m][2] += w * q2;
            }
        }
        __syncthreads();
    }
    const int gx = t0[0] + cxl, gy = t0[1] + cyl;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int gz = t0[2] + czl0 + 2 * m;
        const int64_t off = ((int64_t)gz * g.G[1] + gy) * g.Gxp + gx;
#pragma unroll
        for (int d = 0; d < 3; d++) grid[d * g.cstride + off] = acc[m][d];
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
__global__ void fourier_kernel(FourierArgs<CT> a, CT *__restrict__ grid, CT *__restrict__ grid2) {
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
    CT *c1 = grid + 2 * idx;   // complex (re, im) of component 0
    CT *c2 = grid2 + 2 * idx;
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

template <typename CT>
struct Kahan2 {
    CT s = 0, c = 0;
    __device__ inline void add(CT x) {
        CT y = x - c;
        CT t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

template <typename CT>
__global__ __launch_bounds__(256) void interp_kernel(LRGeom g, int64_t Np, const int32_t *__restrict__ order,
                                                     const double *__restrict__ nodes, CT beta,
                                                     const CT *__restrict__ grid, double *__restrict__ out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Np) return;
    const int64_t i = order ? order[k] : k;
    const int M = g.M, W = 2 * M;
    CT w[3][2 * MAXM];
    int i0[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double u = nodes[d * Np + i] * g.scale[d];
        i0[d] = (int)floor(u) - M + 1;
        for (int q = 0; q < W; q++) w[d][q] = es_kernel<CT>((CT)(((double)(i0[d] + q) - u) / (double)M), beta);
    }
    int ixs[2 * MAXM];
    for (int q = 0; q < W; q++) ixs[q] = wrapi(i0[0] + q, g.G[0]);
    Kahan2<CT> acc[3];
    for (int c = 0; c < W; c++) {
        const int64_t iz = wrapi(i0[2] + c, g.G[2]);
        for (int b = 0; b < W; b++) {
            const int64_t iy = wrapi(i0[1] + b, g.G[1]);
            const int64_t row = (iz * g.G[1] + iy) * g.Gxp;
            const CT wyz = w[2][c] * w[1][b];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const CT *r = grid + d * g.cstride + row;
                CT sx = 0;
                for (int q = 0; q < W; q++) sx += w[0][q] * r[ixs[q]];
                acc[d].add(wyz * sx);
            }
        }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) out[d * Np + i] = (double)acc[d].s;
}

// Type-2 interpolation for half-support m = 4 (8³ stencil): one wavefront per
// node. Lane (a, b) = (lane % 8, lane / 8) owns grid row (y = i0y + b) at
// x = i0x + a and walks the 8 z-planes, so each load instruction touches 8 rows
// of 8 consecutive values; the 64 partial sums are combined by a butterfly
// reduction (pairwise summation, error O(log₂ 512)·ε per node).
template <typename CT>
__global__ __launch_bounds__(256) void interp_kernel_w8(LRGeom g, int64_t Np, const double *__restrict__ nodes,
                                                        CT beta, const CT *__restrict__ grid,
                                                        double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int a = lane & 7, b = lane >> 3;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < Np; i += nwaves) {
        const double ux = nodes[i] * g.scale[0];
        const double uy = nodes[Np + i] * g.scale[1];
        const double uz = nodes[2 * Np + i] * g.scale[2];
        const int i0x = (int)floor(ux) - 3, i0y = (int)floor(uy) - 3, i0z = (int)floor(uz) - 3;
        const CT wx = es_kernel<CT>((CT)(((double)(i0x + a) - ux) / 4.0), beta);
        const CT wy = es_kernel<CT>((CT)(((double)(i0y + b) - uy) / 4.0), beta);
        // Each lane evaluates all 8 z-weights itself. Broadcasting them from
        // lanes 0..7 with a wave shuffle was measured to lose precision in the
        // FP64 build (~1e-7 relative against the oracle); the redundant
        // evaluation is exact and cheap next to the 24 grid loads per lane.
        CT wz[8];
#pragma unroll
        for (int c = 0; c < 8; c++) wz[c] = es_kernel<CT>((CT)(((double)(i0z + c) - uz) / 4.0), beta);
        const int64_t gx = wrapi(i0x + a, g.G[0]);
        const int64_t gy = wrapi(i0y + b, g.G[1]);
        const CT wxy = wx * wy;
        CT acc0 = 0, acc1 = 0, acc2 = 0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int64_t gz = wrapi(i0z + c, g.G[2]);
            const int64_t off = (gz * g.G[1] + gy) * g.Gxp + gx;
            const CT w = wxy * wz[c];
            acc0 += w * grid[off];
            acc1 += w * grid[g.cstride + off];
            acc2 += w * grid[2 * g.cstride + off];
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
    gather_points<CT><<<nblocks(Nq, 256), 256, 0, s>>>(Nq, c->lr_perm, c->points, c->charges, c->lr_pos,
                                                        (CT *)c->lr_chg);
    stage_end(c, ST_LRSORT, s);

    CT *grid = (CT *)c->grid;
    CT *grid2 = (CT *)c->grid2;
    stage_begin(c, ST_SPREAD, s);
    spread_kernel<CT><<<(unsigned)c->ntiles, SPREAD_THREADS, 0, s>>>(g, Nq, c->lr_start, c->lr_pos,
                                                                      (const CT *)c->lr_chg, beta, grid);
    stage_end(c, ST_SPREAD, s);
    VPBS_HIP(c, hipGetLastError());

    stage_begin(c, ST_FFTF, s);
    void *bufs[1] = {grid};
    if (rocfft_execute(c->plan_fwd, bufs, nullptr, c->info_fwd) != rocfft_status_success)
        return fail(c, VPBS_ERR_FFT, "rocfft_execute (forward) failed");
    stage_end(c, ST_FFTF, s);

    FourierArgs<CT> fa;
    for (int d = 0; d < 3; d++) {
        fa.G[d] = c->G[d];
        fa.kmax[d] = c->kmax_mode[d];
        fa.kscale[d] = 2 * M_PI / c->p.Ls[d];
    }
    fa.Hx = c->G[0] / 2 + 1;
    fa.cstride = c->comp_stride / 2;
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
    fourier_kernel<CT><<<nblocks(nspec, 256), 256, 0, s>>>(fa, grid, psi ? grid2 : grid);
    stage_end(c, ST_FOURIER, s);
    VPBS_HIP(c, hipGetLastError());

    stage_begin(c, ST_FFTB, s);
    if (vel) {
        void *b[1] = {grid};
        if (rocfft_execute(c->plan_bwd, b, nullptr, c->info_bwd) != rocfft_status_success)
            return fail(c, VPBS_ERR_FFT, "rocfft_execute (backward) failed");
    }
    if (psi) {
        void *b[1] = {grid2};
        if (rocfft_execute(c->plan_bwd, b, nullptr, c->info_bwd) != rocfft_status_success)
            return fail(c, VPBS_ERR_FFT, "rocfft_execute (backward, psi) failed");
    }
    stage_end(c, ST_FFTB, s);

    stage_begin(c, ST_INTERP, s);
    const int32_t *order = nullptr;  // filament order is already spatially local
    if (g.M == 4) {
        const unsigned nb = (unsigned)std::min<int64_t>((Np + 3) / 4, 1 << 20);  // 4 nodes (waves) per block
        if (vel) interp_kernel_w8<CT><<<nb, 256, 0, s>>>(g, Np, c->nodes, beta, grid, c->v_lr);
        if (psi) interp_kernel_w8<CT><<<nb, 256, 0, s>>>(g, Np, c->nodes, beta, grid2, c->psi_lr);
    } else {
        if (vel) interp_kernel<CT><<<nblocks(Np, 256), 256, 0, s>>>(g, Np, order, c->nodes, beta, grid, c->v_lr);
        if (psi) interp_kernel<CT><<<nblocks(Np, 256), 256, 0, s>>>(g, Np, order, c->nodes, beta, grid2, c->psi_lr);
    }
    stage_end(c, ST_INTERP, s);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

bool rocfft_ok(rocfft_status st) { return st == rocfft_status_success; }

}  // namespace

double es_beta(int m, double sigma) { return 0.98 * M_PI * m * (2.0 - 1.0 / sigma); }

int oversampled_size(int N, double sigma) {
    // smallest multiple of 16 with no prime factor above 5, ≥ ⌈σN⌉
    int n = (int)std::ceil(sigma * N - 1e-9);
    for (;; n++) {
        if (n % 16) continue;
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
        if (c->G[d] < 4 * m) c->G[d] = ((4 * m + 15) / 16) * 16;
        c->ntile[d] = c->G[d] / (d == 0 ? TX : (d == 1 ? TY : TZ));
    }
    c->ntiles = (int64_t)c->ntile[0] * c->ntile[1] * c->ntile[2];
    c->Gxp = 2 * (c->G[0] / 2 + 1);
    c->comp_stride = c->Gxp * c->G[1] * c->G[2];
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

    // grid buffers
    const size_t es = c->p.precision == VPBS_F32 ? sizeof(float) : sizeof(double);
    VPBS_HIP(c, hipMalloc(&c->grid, es * 3 * (size_t)c->comp_stride));

    // rocFFT plans: 3 transforms (components) of G0×G1×G2, x fastest, in place
    rocfft_precision prec = c->p.precision == VPBS_F32 ? rocfft_precision_single : rocfft_precision_double;
    size_t len[3] = {(size_t)c->G[0], (size_t)c->G[1], (size_t)c->G[2]};
    const size_t Hx = c->G[0] / 2 + 1;
    size_t rstr[3] = {1, (size_t)c->Gxp, (size_t)c->Gxp * c->G[1]};
    size_t cstr[3] = {1, Hx, Hx * c->G[1]};
    size_t rdist = (size_t)c->comp_stride, cdist = (size_t)c->comp_stride / 2;
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
    if (!rocfft_ok(rocfft_plan_create(&c->plan_fwd, rocfft_placement_inplace, rocfft_transform_type_real_forward,
                                      prec, 3, len, 3, df)) ||
        !rocfft_ok(rocfft_plan_create(&c->plan_bwd, rocfft_placement_inplace, rocfft_transform_type_real_inverse,
                                      prec, 3, len, 3, db)))
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
    if (psi && !c->grid2) VPBS_HIP(c, hipMalloc(&c->grid2, es * 3 * (size_t)c->comp_stride));
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
    if (c->grid2) (void)hipFree(c->grid2);
    if (c->deconv) (void)hipFree(c->deconv);
    c->plan_fwd = c->plan_bwd = nullptr;
    c->info_fwd = c->info_bwd = nullptr;
    c->fft_work = c->grid = c->grid2 = nullptr;
    c->deconv = nullptr;
}

int merge_outputs(vpbs_ctx *c, const double *a, const double *b, double *o, int64_t n, hipStream_t s) {
    add3<<<nblocks(n, 256), 256, 0, s>>>(n, a, b, o);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

}  // namespace vpbs
