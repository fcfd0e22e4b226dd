// Short-range part of the Ewald-split Biot–Savart sum on gfx950.
//
//   local_term_kernel   LIA term, FP64 (shortrange/local_integrals.jl:106-165)
//   cell keys + sort    device cell list (CellLists.jl:131-137,194-204), built by a
//                       deterministic counting sort instead of an atomic linked list
//   pair_wave_kernel    erfc/Gaussian pair sum over the (2M+1)³ neighbour cells,
//                       one wavefront per target with compaction of the accepted pairs
//                       (shortrange.jl:216-273, integrands.jl:2-18, gaussian.jl:273-279)
//   self_kernel         removal of the local-segment contribution (local_integrals.jl:14-95)
//
// Precision. CT = double is the plain FP64 path. CT = float is the FP32-bulk
// path: node/source coordinates are differenced and folded in FP64 and the
// cut-off test r² ≤ rcut² is taken in FP64 (so both paths accept exactly the
// same pairs); the weights, cross products and the Kahan-compensated per-node
// sums are FP32; local segments are excluded from the pair sum and their smooth
// long-range part (erf form) is removed in FP64 (avoid_explicit_erf = false
// semantics, params.jl:289-298), which avoids the FP32 cancellation of the
// singular local integrals.
#include "vpbs_internal.hpp"
#include "device_util.hpp"

#include <cmath>

namespace vpbs {

namespace {

constexpr double TWO_OVER_SQRT_PI = 1.1283791670955126;


struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ inline void add(float x) {
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// ----------------------------------------------------------------------------
// LIA term (FP64 in both paths). Replaces outputs.

__global__ void local_term_kernel(int64_t Np, const double *__restrict__ sp, const double *__restrict__ spp,
                                  const double *__restrict__ dlm, const double *__restrict__ dlp, double pref,
                                  double cut_v, double cut_psi, int do_vel, int do_psi, double *__restrict__ v,
                                  double *__restrict__ psi) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    double delta = sqrt(dlm[i] * dlp[i]);
    double s1x = sp[i], s1y = sp[Np + i], s1z = sp[2 * Np + i];
    double mu2 = dist2_nocontract(s1x, s1y, s1z);
    double muinv = 1.0 / sqrt(mu2);
    if (do_vel) {
        double s2x = spp[i], s2y = spp[Np + i], s2z = spp[2 * Np + i];
        double m3 = muinv * muinv * muinv;
        double lg = log(delta / cut_v);
        v[i] = pref * (lg * ((s1y * s2z - s1z * s2y) * m3));
        v[Np + i] = pref * (lg * ((s1z * s2x - s1x * s2z) * m3));
        v[2 * Np + i] = pref * (lg * ((s1x * s2y - s1y * s2x) * m3));
    }
    if (do_psi) {
        double lg = 2 * log(delta / cut_psi);
        psi[i] = pref * (lg * (s1x * muinv));
        psi[Np + i] = pref * (lg * (s1y * muinv));
        psi[2 * Np + i] = pref * (lg * (s1z * muinv));
    }
}

__global__ void zero_f64(double *__restrict__ a, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// ----------------------------------------------------------------------------
// Cell list

// key[j] = linear cell of point j; if cell != nullptr also cell[j] = (ix, iy, iz, key), so that the
// pair kernel reads its target's cell instead of recomputing three FP64 divisions per node.
__global__ void cell_keys(int64_t n, const double *__restrict__ pos, Box box, int3 Nc, int32_t *__restrict__ key,
                          int4 *__restrict__ cell) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int ix = cell_index(pos[j], box.L[0], Nc.x);
    int iy = cell_index(pos[n + j], box.L[1], Nc.y);
    int iz = cell_index(pos[2 * n + j], box.L[2], Nc.z);
    // Points folded into [0, L) always give indices in [0, N); clamp defensively.
    ix = min(max(ix, 0), Nc.x - 1);
    iy = min(max(iy, 0), Nc.y - 1);
    iz = min(max(iz, 0), Nc.z - 1);
    const int k = (iz * Nc.y + iy) * Nc.x + ix;
    key[j] = k;
    if (cell) cell[j] = make_int4(ix, iy, iz, k);
}

template <typename CT>
__global__ void gather_sources(int64_t Nq, const int32_t *__restrict__ perm, const double *__restrict__ pts,
                               const double *__restrict__ chg, double *__restrict__ spos, CT *__restrict__ schg,
                               int32_t *__restrict__ sidx) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Nq) return;
    int32_t j = perm[s];
    sidx[s] = j;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        spos[d * Nq + s] = pts[d * Nq + j];
        schg[d * Nq + s] = (CT)chg[d * Nq + j];
    }
}

// ----------------------------------------------------------------------------
// Pair kernel

template <typename CT>
struct PairArgs {
    int64_t Np, Nq;
    const int32_t *tperm;
    const int4 *tcell;  // per node: (ix, iy, iz, key) of its cell
    const double *nodes;
    const int32_t *cstart;
    const double *spos;
    const CT *schg;
    const int32_t *sidx;
    const int32_t *idx_prev;
    int nquad;
    Box box;
    int3 Nc;
    int M;
    double cinv[3];  // Nc / L: cells per unit length
    float cinvf[3], cszf[3];  // the same in FP32, and L / Nc (the rows' cut to the rcut sphere)
    float rcut2f;
    double rcut2;
    CT alpha;
    double pref;
    int include_local;
    double *v;
    double *psi;
    int64_t *counts;
    uint64_t *hashes;
    int32_t *cands;  // stats: candidate sources tested per node
};


template <typename CT>
struct Acc3;
template <>
struct Acc3<double> {
    double a[3] = {0, 0, 0};
    __device__ inline void add(int d, double x) { a[d] += x; }
    __device__ inline double get(int d) const { return a[d]; }
};
template <>
struct Acc3<float> {
    Kahan k[3];
    __device__ inline void add(int d, float x) { k[d].add(x); }
    __device__ inline double get(int d) const { return (double)k[d].s; }
};

// FP64: correctly rounded operations (1e-12 parity with the oracle). FP32-bulk: the
// hardware reciprocal square root (r = r²·(1/√r²)) and base-2 exponential, a few ulp
// each, well inside the FP32 error budget and ~25 instructions cheaper per pair than a
// correctly rounded square root and division.
__device__ inline double erfc_t(double x) { return erfc(x); }
__device__ inline float erfc_t(float x) { return erfcf(x); }
__device__ inline double exp_t(double x) { return exp(x); }
__device__ inline float exp_t(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896340736f); }
// 1/r and r from r²
__device__ inline void rinv_r(double r2, double &rinv, double &r) {
    r = sqrt(r2);
    rinv = 1.0 / r;
}
__device__ inline void rinv_r(float r2, float &rinv, float &r) {
    rinv = __builtin_amdgcn_rsqf(r2);
    r = r2 * rinv;
}


// One wavefront per target node (cell-sorted order; XCD-aware block order):
//  1) lanes build the source runs of the (2M+1)² neighbour cell rows (≤ 2 contiguous
//     runs per row with the periodic wrap along x), each cut to the cells that the rcut
//     sphere around the node can reach, and a wave scan gives each row's candidate offset;
//  2) the candidates are tested 64 at a time, consecutive lanes reading consecutive
//     sources: FP64 separation, minimum-image fold and r² ≤ rcut² exactly as in the
//     reference (shortrange.jl:255-256), local segments excluded when required; the
//     accepted pairs are appended in candidate order to an LDS queue (ballot compaction);
//  3) each time 64 are queued (and at the end) every lane evaluates one queued pair
//     (erfc/Gaussian weights, cross product; integrands.jl:2-18, gaussian.jl:273-279)
//     into its own compensated accumulator, so the expensive part runs on full waves;
//  4) the 64 partial sums are combined by a fixed butterfly in FP64.
// A candidate's row: every row meeting a batch marks its first candidate in it and a
// max-scan over the lanes carries the marks forward. The cut drops only sources beyond
// rcut, so the accepted pair set is the thread-per-target one, in the same order; the
// summation order is fixed (deterministic) but differs from a sequential sum.
constexpr int PAIR_MAXROWS = 81;  // (2M+1)² for M ≤ 4

template <typename CT, bool VEL, bool PSI, bool STATS>
__global__ __launch_bounds__(64, 4) void pair_wave_kernel(PairArgs<CT> a) {
    // per row: (candidates before it, run 0 start, run 0 length, run 1 start - run 0 length)
    __shared__ int4 r_run[PAIR_MAXROWS];
    __shared__ int q_s[128];
    __shared__ CT q_r[4][128];  // rx, ry, rz, r² of the queued pairs
    __shared__ int s_mark[64];  // rows meeting the fetched batch, at their first candidate in it

    const int lane = threadIdx.x;
    const int64_t Np = a.Np, Nq = a.Nq;
    const int64_t per = ((int64_t)gridDim.x + 7) / 8;
    const int64_t k = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (k >= Np) return;
    const int32_t i = a.tperm[k];
    const double x = a.nodes[i], y = a.nodes[Np + i], z = a.nodes[2 * Np + i];
    const int4 tc = a.tcell[i];  // the node's cell, as computed for the sort
    const int cx = tc.x, cy = tc.y, cz = tc.z;
    // first quadrature points of the node's right and left segments (source indices are int32)
    const int jr0 = a.nquad * i, jl0 = a.nquad * a.idx_prev[i];
    const int M = a.M, W2 = 2 * M + 1, nrows = W2 * W2;
    // The minimum-image fold changes no separation of a node whose neighbour cells do not
    // wrap, when every direction has at least 2M + 2 cells: then |x - s| < (M + 1)·L/Nc ≤ L/2
    // for every candidate, up to rounding at exactly L/2, which r² ≤ rcut² (rcut < L/2)
    // rejects either way. Such nodes (most of them) skip the fold.
    const bool wraps = cx < M || cy < M || cz < M || cx + M >= a.Nc.x || cy + M >= a.Nc.y || cz + M >= a.Nc.z ||
                       min(min(a.Nc.x, a.Nc.y), a.Nc.z) < 2 * M + 2;
    // row r = (dz, dy) = (r / W2 - M, r % W2 - M); the quotient from a float product with the
    // hardware reciprocal (1 ulp) is exact: (r + 1/2) / W2 stays ≥ 1/18 away from an integer
    // for r < 128, W2 ≤ 9
    const float rcp_w2 = __builtin_amdgcn_rcpf((float)W2);

    // 1) lanes take rows lane and lane + 64. Each row keeps only the cells that can hold a
    //    source within rcut: rows whose (y, z) gap to the node exceeds rcut are empty, the
    //    others are cut to the x-cells within √(rcut² − gap²) of the node (widened by 1e-4
    //    cells: a conservative cut that drops only sources the r² test rejects, so the
    //    accepted pairs and their order are unchanged)
    // place in the cell (FP64, then FP32: the cut below is widened by 1e-4 cells, far above
    // the FP32 rounding of these O(M)-cell quantities)
    const float u[3] = {(float)(x * a.cinv[0] - cx), (float)(y * a.cinv[1] - cy), (float)(z * a.cinv[2] - cz)};
    int len[2] = {0, 0};
    int4 run[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int r = lane + 64 * h;
        if (r < nrows) {
            const int qz = (int)(((float)r + 0.5f) * rcp_w2);
            const int dz = qz - M, dy = (r - qz * W2) - M;
            int zz = cz + dz, yy = cy + dy;
            zz += (zz < 0) ? a.Nc.z : 0;
            zz -= (zz >= a.Nc.z) ? a.Nc.z : 0;
            yy += (yy < 0) ? a.Nc.y : 0;
            yy -= (yy >= a.Nc.y) ? a.Nc.y : 0;
            const int64_t row = ((int64_t)zz * a.Nc.y + yy) * a.Nc.x;
            const float gy = fmaxf((dy > 0 ? dy - u[1] : (dy < 0 ? u[1] - (dy + 1) : 0.f)) - 1e-4f, 0.f) * a.cszf[1];
            const float gz = fmaxf((dz > 0 ? dz - u[2] : (dz < 0 ? u[2] - (dz + 1) : 0.f)) - 1e-4f, 0.f) * a.cszf[2];
            const float rem = a.rcut2f - (gy * gy + gz * gz);
            int lo = 1, hi = 0;  // empty
            if (rem >= 0.f) {
                const float ext = sqrtf(rem) * a.cinvf[0] + 1e-4f;
                lo = cx + max((int)floorf(u[0] - ext), -M);
                hi = cx + min((int)floorf(u[0] + ext), M);
            }
            int ca0 = 0, cb0 = -1, ca1 = 0, cb1 = -1;
            if (lo > hi) {
            } else if (hi < 0) {
                ca0 = lo + a.Nc.x, cb0 = hi + a.Nc.x;
            } else if (lo < 0) {
                ca0 = lo + a.Nc.x, cb0 = a.Nc.x - 1, cb1 = hi;
            } else if (lo >= a.Nc.x) {
                ca0 = lo - a.Nc.x, cb0 = hi - a.Nc.x;
            } else if (hi >= a.Nc.x) {
                ca0 = lo, cb0 = a.Nc.x - 1, cb1 = hi - a.Nc.x;
            } else {
                ca0 = lo, cb0 = hi;
            }
            int s0 = 0, n0 = 0, s1 = 0, n1 = 0;
            if (cb0 >= ca0) {
                s0 = a.cstart[row + ca0];
                n0 = a.cstart[row + cb0 + 1] - s0;
            }
            if (cb1 >= ca1) {
                s1 = a.cstart[row + ca1];
                n1 = a.cstart[row + cb1 + 1] - s1;
            }
            run[h] = make_int4(0, s0, n0, s1 - n0);
            len[h] = n0 + n1;
        }
    }
    // inclusive prefix: lane r holds the candidates before row r + 1 (incl0) or row r + 65 (incl1)
    const int incl0 = wave_incl_scan(len[0]);
    const int tot0 = __builtin_amdgcn_readlane(incl0, 63);
    const int incl1 = nrows > 64 ? tot0 + wave_incl_scan(len[1]) : tot0;
    if (lane < nrows) {
        run[0].x = incl0 - len[0];
        r_run[lane] = run[0];
    }
    if (lane + 64 < nrows) {
        run[1].x = incl1 - len[1];
        r_run[lane + 64] = run[1];
    }
    const int ncand = nrows > 64 ? __builtin_amdgcn_readlane(incl1, nrows - 65) : __builtin_amdgcn_readlane(incl0, nrows - 1);
    __syncthreads();
    Acc3<CT> av, ap;
    int64_t cnt = 0;
    uint64_t hsum = 0;
    const CT two_over_sqrt_pi = (CT)TWO_OVER_SQRT_PI;
    auto eval = [&](int e) {
        const int s = q_s[e];
        const CT crx = q_r[0][e], cry = q_r[1][e], crz = q_r[2][e];
        CT rinv, r;
        rinv_r(q_r[3][e], rinv, r);
        const CT ar = a.alpha * r;
        const CT wa = erfc_t(ar);
        const CT qx = a.schg[s], qy = a.schg[Nq + s], qz = a.schg[2 * Nq + s];
        if (VEL) {
            const CT wb = two_over_sqrt_pi * ar * exp_t(-(ar * ar));
            const CT f = (wa + wb) * (rinv * rinv * rinv);
            av.add(0, f * (qy * crz - qz * cry));
            av.add(1, f * (qz * crx - qx * crz));
            av.add(2, f * (qx * cry - qy * crx));
        }
        if (PSI) {
            const CT f = wa * rinv;
            ap.add(0, f * qx);
            ap.add(1, f * qy);
            ap.add(2, f * qz);
        }
    };

    int nqd = 0;   // queued pairs (uniform)
    // Locates candidate b + lane and issues the loads of its source, index and position
    // together, so that the loads of one batch stay in flight while the previous batch is
    // tested. Its row: every non-empty row that meets the batch marks its first candidate in
    // it, and a max-scan carries the marks forward. Lanes past the last candidate read source
    // 0 (there is at least one source when there are candidates).
    auto fetch = [&](int b, int &fs, int32_t &fj, double &fx, double &fy, double &fz) {
        s_mark[lane] = -1;
        wave_lds_order();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int r = lane + 64 * h;
            if (r < nrows && len[h] > 0) {
                const int st = run[h].x, en = st + len[h];
                if (st < b + 64 && en > b) s_mark[st > b ? st - b : 0] = r;
            }
        }
        wave_lds_order();
        const int crow = wave_incl_max(s_mark[lane]);
        const int cnd = b + lane;
        fs = 0;
        if (cnd < ncand) {
            const int4 rr = r_run[crow];
            const int o = cnd - rr.x;
            fs = (o < rr.z ? rr.y : rr.w) + o;
        }
        fj = a.sidx[fs];
        fx = a.spos[fs];
        fy = a.spos[Nq + fs];
        fz = a.spos[2 * Nq + fs];
    };
    // current batch's source (s, j, position); the next batch's is fetched before the
    // current one is tested, so its load latency overlaps the test and the evaluation
    int s = 0;
    int32_t j = 0;
    double sx = 0, sy = 0, sz = 0;
    if (ncand > 0) fetch(0, s, j, sx, sy, sz);
    for (int base = 0; base < ncand; base += 64) {
        int s_n;
        int32_t j_n;
        double sx_n, sy_n, sz_n;
        fetch(base + 64, s_n, j_n, sx_n, sy_n, sz_n);
        // 2) test 64 candidates
        const int cnd = base + lane;
        const bool local = !a.include_local &&
                           ((unsigned)(j - jr0) < (unsigned)a.nquad || (unsigned)(j - jl0) < (unsigned)a.nquad);
        double rx = x - sx, ry = y - sy, rz = z - sz;
        if (wraps) {  // wave-uniform
            rx = fold_sep(rx, a.box.L[0], a.box.Lh[0]);
            ry = fold_sep(ry, a.box.L[1], a.box.Lh[1]);
            rz = fold_sep(rz, a.box.L[2], a.box.Lh[2]);
        }
        const double r2 = dist2_nocontract(rx, ry, rz);
        const bool acc = cnd < ncand && !local && r2 <= a.rcut2;
        if (STATS && acc) {
            cnt++;
            hsum += mix64((uint64_t)j);
        }
        const unsigned long long mask = __ballot(acc);
        if (acc) {
            const int pos = nqd + __popcll(mask & ((1ull << lane) - 1ull));
            q_s[pos] = s;
            q_r[0][pos] = (CT)rx;
            q_r[1][pos] = (CT)ry;
            q_r[2][pos] = (CT)rz;
            q_r[3][pos] = (CT)r2;
        }
        nqd += __popcll(mask);
        __syncthreads();
        // 3) full waves of queued pairs
        if (nqd >= 64) {
            eval(lane);
            const int rest = nqd - 64;  // ≤ 63: move entries [64, nqd) to the front
            int ms = 0;
            CT mr[4];
            if (lane < rest) {
                ms = q_s[64 + lane];
#pragma unroll
                for (int d = 0; d < 4; d++) mr[d] = q_r[d][64 + lane];
            }
            __syncthreads();
            if (lane < rest) {
                q_s[lane] = ms;
#pragma unroll
                for (int d = 0; d < 4; d++) q_r[d][lane] = mr[d];
            }
            nqd = rest;
            __syncthreads();
        }
        s = s_n;
        j = j_n;
        sx = sx_n;
        sy = sy_n;
        sz = sz_n;
    }
    if (lane < nqd) eval(lane);

    // 4) fixed-order reduction of the lane partial sums (DPP steps in FP64; total in lane 63)
    if (VEL) {
        double v3[3];
#pragma unroll
        for (int d = 0; d < 3; d++) v3[d] = wave_sum_f64_last(av.get(d));
        if (lane == 63) {
            a.v[i] += a.pref * v3[0];
            a.v[Np + i] += a.pref * v3[1];
            a.v[2 * Np + i] += a.pref * v3[2];
        }
    }
    if (PSI) {
        double p3[3];
#pragma unroll
        for (int d = 0; d < 3; d++) p3[d] = wave_sum_f64_last(ap.get(d));
        if (lane == 63) {
            a.psi[i] += a.pref * p3[0];
            a.psi[Np + i] += a.pref * p3[1];
            a.psi[2 * Np + i] += a.pref * p3[2];
        }
    }
    if (STATS) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            cnt += __shfl_xor(cnt, off, 64);
            hsum += __shfl_xor(hsum, off, 64);
        }
        if (lane == 0) {
            a.counts[i] = cnt;
            a.hashes[i] = hsum;
            a.cands[i] = ncand;
        }
    }
}

// Tiled pair kernel: one wavefront per tile of PT_T³ cells of the cell list. The sources of
// the union of the tile's neighbour cells ((PT_T + 2M)² rows of PT_T + 2M cells) are staged in
// LDS once (chunks of PT_CAP; a larger union is staged again per target), then the tile's
// targets are taken one after the other, each testing every staged source 64 at a time
// (FP64 separation, fold where the union wraps, r² ≤ rcut², local segments excluded when
// required) with the compaction and evaluation of pair_wave_kernel. Every source within rcut
// of a target lies in its (2M + 1)³ neighbour cells, which the union contains exactly once
// (host: at least PT_T + 2M cells per direction), so the accepted pairs are those of the
// per-target kernel; only the summation order differs (fixed, deterministic). The row setup
// and the global source loads are shared by the tile's targets (~3.6 per tile at 1M nodes).
// Measured at 1M nodes it is slower than pair_wave_kernel (FP32 6.3 against 1.3 ms, FP64 9.1
// against 3.0 ms: 20 KB of LDS per wavefront leave 2 wavefronts per SIMD), so it runs only
// when VPBS_PAIR=tile is set at context creation.
constexpr int PT_T = 2;
constexpr int PT_CAP = 512;
constexpr int PT_MAXROWS = (PT_T + 8) * (PT_T + 8);  // (PT_T + 2M)² for M ≤ 4

template <typename CT, bool VEL, bool PSI, bool STATS>
__global__ __launch_bounds__(64) void pair_tile_kernel(PairArgs<CT> a, const int32_t *__restrict__ tstart, int3 nt) {
    __shared__ int4 r_run[PT_MAXROWS];  // (candidates before the row, run 0 start, run 0 length, run 1 start - run 0 length)
    __shared__ double st_p[3][PT_CAP];
    __shared__ int st_j[PT_CAP], st_s[PT_CAP];
    __shared__ int q_s[128];
    __shared__ CT q_r[4][128];

    const int lane = threadIdx.x;
    const int64_t Np = a.Np, Nq = a.Nq;
    const int64_t ntiles = (int64_t)nt.x * nt.y * nt.z;
    const int64_t per = ((int64_t)gridDim.x + 7) / 8;
    const int64_t k = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (k >= ntiles) return;
    const int3 Nc = a.Nc;
    const int tx = (int)(k % nt.x), ty = (int)((k / nt.x) % nt.y), tz = (int)(k / ((int64_t)nt.x * nt.y));
    const int x0 = tx * PT_T, x1 = min(x0 + PT_T, Nc.x);
    const int y0 = ty * PT_T, y1 = min(y0 + PT_T, Nc.y);
    const int z0 = tz * PT_T, z1 = min(z0 + PT_T, Nc.z);

    // the tile's targets: one run of the cell-sorted targets per (z, y) row of the tile
    int t_beg = 0, t_len = 0;
    if (lane < PT_T * PT_T) {
        const int yy = y0 + lane % PT_T, zz = z0 + lane / PT_T;
        if (yy < y1 && zz < z1) {
            const int64_t row = ((int64_t)zz * Nc.y + yy) * Nc.x;
            t_beg = tstart[row + x0];
            t_len = tstart[row + x1] - t_beg;
        }
    }
    const int t_incl = wave_incl_scan(t_len);
    const int ntgt = __builtin_amdgcn_readlane(t_incl, PT_T * PT_T - 1);
    if (ntgt == 0) return;

    // union rows (z, y) of x-cells [x0 - M, x1 - 1 + M], at most two runs with the periodic wrap
    const int M = a.M;
    const int ux0 = x0 - M, ux1 = x1 - 1 + M;
    const int nry = (y1 - y0) + 2 * M, nrz = (z1 - z0) + 2 * M, nrows = nry * nrz;
    int ca0, cb0, ca1 = 0, cb1 = -1;
    if (ux0 < 0) {
        ca0 = ux0 + Nc.x, cb0 = Nc.x - 1, cb1 = ux1;
    } else if (ux1 >= Nc.x) {
        ca0 = ux0, cb0 = Nc.x - 1, cb1 = ux1 - Nc.x;
    } else {
        ca0 = ux0, cb0 = ux1;
    }
    int len[2] = {0, 0};
    int4 run[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int r = lane + 64 * h;
        if (r < nrows) {
            const int qz = r / nry;
            int zz = z0 - M + qz, yy = y0 - M + (r - qz * nry);
            zz += (zz < 0) ? Nc.z : 0;
            zz -= (zz >= Nc.z) ? Nc.z : 0;
            yy += (yy < 0) ? Nc.y : 0;
            yy -= (yy >= Nc.y) ? Nc.y : 0;
            const int64_t row = ((int64_t)zz * Nc.y + yy) * Nc.x;
            const int s0 = a.cstart[row + ca0], n0 = a.cstart[row + cb0 + 1] - s0;
            int s1 = 0, n1 = 0;
            if (cb1 >= ca1) {
                s1 = a.cstart[row + ca1];
                n1 = a.cstart[row + cb1 + 1] - s1;
            }
            run[h] = make_int4(0, s0, n0, s1 - n0);
            len[h] = n0 + n1;
        }
    }
    const int incl0 = wave_incl_scan(len[0]);
    const int tot0 = __builtin_amdgcn_readlane(incl0, 63);
    const int incl1 = nrows > 64 ? tot0 + wave_incl_scan(len[1]) : tot0;
    if (lane < nrows) {
        run[0].x = incl0 - len[0];
        r_run[lane] = run[0];
    }
    if (lane + 64 < nrows) {
        run[1].x = incl1 - len[1];
        r_run[lane + 64] = run[1];
    }
    const int nunion = nrows > 64 ? __builtin_amdgcn_readlane(incl1, nrows - 65)
                                  : __builtin_amdgcn_readlane(incl0, nrows - 1);
    __syncthreads();
    // Separations need the minimum-image fold only where the union wraps, or when a direction
    // has fewer than 2(PT_T + M) cells; otherwise |x - s| < (PT_T + M)·L/Nc ≤ L/2.
    const bool wraps = ux0 < 0 || ux1 >= Nc.x || y0 - M < 0 || y1 - 1 + M >= Nc.y || z0 - M < 0 ||
                       z1 - 1 + M >= Nc.z || min(min(Nc.x, Nc.y), Nc.z) < 2 * (PT_T + M);

    // stages union candidates [c0, c0 + n) into LDS slots [0, n)
    auto stage = [&](int c0, int n) {
        for (int e = lane; e < n; e += 64) {
            const int cnd = c0 + e;
            int lo = 0, hi = nrows - 1;  // last row with candidates before it ≤ cnd
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (r_run[mid].x <= cnd) lo = mid; else hi = mid - 1;
            }
            const int4 rr = r_run[lo];
            const int o = cnd - rr.x;
            const int sidx = (o < rr.z ? rr.y : rr.w) + o;
            st_p[0][e] = a.spos[sidx];
            st_p[1][e] = a.spos[Nq + sidx];
            st_p[2][e] = a.spos[2 * Nq + sidx];
            st_j[e] = a.sidx[sidx];
            st_s[e] = sidx;
        }
    };
    const int nchunks = (nunion + PT_CAP - 1) / PT_CAP;
    if (nchunks == 1) stage(0, nunion);
    __syncthreads();

    const CT two_over_sqrt_pi = (CT)TWO_OVER_SQRT_PI;
    int run_r = 0;  // uniform: target run of the current target
    for (int u = 0; u < ntgt; u++) {
        while (u >= __builtin_amdgcn_readlane(t_incl, run_r)) run_r++;
        const int slot = __builtin_amdgcn_readlane(t_beg, run_r) + u -
                         (__builtin_amdgcn_readlane(t_incl, run_r) - __builtin_amdgcn_readlane(t_len, run_r));
        const int32_t i = a.tperm[slot];
        const double x = a.nodes[i], y = a.nodes[Np + i], z = a.nodes[2 * Np + i];
        const int jr0 = a.nquad * i, jl0 = a.nquad * a.idx_prev[i];
        Acc3<CT> av, ap;
        int64_t cnt = 0;
        uint64_t hsum = 0;
        auto eval = [&](int e) {
            const int s = q_s[e];
            const CT crx = q_r[0][e], cry = q_r[1][e], crz = q_r[2][e];
            CT rinv, r;
            rinv_r(q_r[3][e], rinv, r);
            const CT ar = a.alpha * r;
            const CT wa = erfc_t(ar);
            const CT qx = a.schg[s], qy = a.schg[Nq + s], qz = a.schg[2 * Nq + s];
            if (VEL) {
                const CT wb = two_over_sqrt_pi * ar * exp_t(-(ar * ar));
                const CT f = (wa + wb) * (rinv * rinv * rinv);
                av.add(0, f * (qy * crz - qz * cry));
                av.add(1, f * (qz * crx - qx * crz));
                av.add(2, f * (qx * cry - qy * crx));
            }
            if (PSI) {
                const CT f = wa * rinv;
                ap.add(0, f * qx);
                ap.add(1, f * qy);
                ap.add(2, f * qz);
            }
        };
        int nqd = 0;
        for (int c0 = 0; c0 < nunion; c0 += PT_CAP) {
            const int n = min(PT_CAP, nunion - c0);
            if (nchunks > 1) {
                __syncthreads();
                stage(c0, n);
                __syncthreads();
            }
            for (int b = 0; b < n; b += 64) {
                const int e = b + lane;
                const bool valid = e < n;
                const int ee = valid ? e : 0;
                const double sx = st_p[0][ee], sy = st_p[1][ee], sz = st_p[2][ee];
                const int32_t j = st_j[ee];
                const bool local = !a.include_local && ((unsigned)(j - jr0) < (unsigned)a.nquad ||
                                                        (unsigned)(j - jl0) < (unsigned)a.nquad);
                double rx = x - sx, ry = y - sy, rz = z - sz;
                if (wraps) {  // wave-uniform
                    rx = fold_sep(rx, a.box.L[0], a.box.Lh[0]);
                    ry = fold_sep(ry, a.box.L[1], a.box.Lh[1]);
                    rz = fold_sep(rz, a.box.L[2], a.box.Lh[2]);
                }
                const double r2 = dist2_nocontract(rx, ry, rz);
                const bool acc = valid && !local && r2 <= a.rcut2;
                if (STATS && acc) {
                    cnt++;
                    hsum += mix64((uint64_t)j);
                }
                const unsigned long long mask = __ballot(acc);
                if (acc) {
                    const int pos = nqd + __popcll(mask & ((1ull << lane) - 1ull));
                    q_s[pos] = st_s[ee];
                    q_r[0][pos] = (CT)rx;
                    q_r[1][pos] = (CT)ry;
                    q_r[2][pos] = (CT)rz;
                    q_r[3][pos] = (CT)r2;
                }
                nqd += __popcll(mask);
                __syncthreads();
                if (nqd >= 64) {
                    eval(lane);
                    const int rest = nqd - 64;
                    int ms = 0;
                    CT mr[4];
                    if (lane < rest) {
                        ms = q_s[64 + lane];
#pragma unroll
                        for (int d = 0; d < 4; d++) mr[d] = q_r[d][64 + lane];
                    }
                    __syncthreads();
                    if (lane < rest) {
                        q_s[lane] = ms;
#pragma unroll
                        for (int d = 0; d < 4; d++) q_r[d][lane] = mr[d];
                    }
                    nqd = rest;
                    __syncthreads();
                }
            }
        }
        if (lane < nqd) eval(lane);
        __syncthreads();  // the queue is reused by the next target
        if (VEL) {
            double v3[3] = {av.get(0), av.get(1), av.get(2)};
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
#pragma unroll
                for (int d = 0; d < 3; d++) v3[d] += __shfl_xor(v3[d], off, 64);
            if (lane == 0) {
                a.v[i] += a.pref * v3[0];
                a.v[Np + i] += a.pref * v3[1];
                a.v[2 * Np + i] += a.pref * v3[2];
            }
        }
        if (PSI) {
            double p3[3] = {ap.get(0), ap.get(1), ap.get(2)};
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
#pragma unroll
                for (int d = 0; d < 3; d++) p3[d] += __shfl_xor(p3[d], off, 64);
            if (lane == 0) {
                a.psi[i] += a.pref * p3[0];
                a.psi[Np + i] += a.pref * p3[1];
                a.psi[2 * Np + i] += a.pref * p3[2];
            }
        }
        if (STATS) {
            // candidates of the node's own (2M + 1)³ neighbour cells (the per-target work measure)
            const int4 tc = a.tcell[i];
            const int W2 = 2 * M + 1;
            int nc = 0;
            for (int r = lane; r < W2 * W2; r += 64) {
                int zz = tc.z + r / W2 - M, yy = tc.y + r % W2 - M;
                zz += (zz < 0) ? Nc.z : 0;
                zz -= (zz >= Nc.z) ? Nc.z : 0;
                yy += (yy < 0) ? Nc.y : 0;
                yy -= (yy >= Nc.y) ? Nc.y : 0;
                const int64_t row = ((int64_t)zz * Nc.y + yy) * Nc.x;
                const int lo = tc.x - M, hi = tc.x + M;
                if (lo < 0)
                    nc += (a.cstart[row + Nc.x] - a.cstart[row + lo + Nc.x]) + (a.cstart[row + hi + 1] - a.cstart[row]);
                else if (hi >= Nc.x)
                    nc += (a.cstart[row + Nc.x] - a.cstart[row + lo]) + (a.cstart[row + hi - Nc.x + 1] - a.cstart[row]);
                else
                    nc += a.cstart[row + hi + 1] - a.cstart[row + lo];
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                cnt += __shfl_xor(cnt, off, 64);
                hsum += __shfl_xor(hsum, off, 64);
                nc += __shfl_xor(nc, off, 64);
            }
            if (lane == 0) {
                a.counts[i] = cnt;
                a.hashes[i] = hsum;
                a.cands[i] = nc;
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Self-interaction removal (FP64 in both paths)

__global__ void self_kernel(int64_t Np, int64_t Nq, const double *__restrict__ nodes,
                            const int32_t *__restrict__ idx_prev, const double *__restrict__ pts,
                            const double *__restrict__ chg, int nquad, Box box, double alpha, double pref,
                            int full, int do_vel, int do_psi, double *__restrict__ v, double *__restrict__ psi) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    const double x = nodes[i], y = nodes[Np + i], z = nodes[2 * Np + i];
    double vx = v ? v[i] : 0, vy = v ? v[Np + i] : 0, vz = v ? v[2 * Np + i] : 0;
    double px = psi ? psi[i] : 0, py = psi ? psi[Np + i] : 0, pz = psi ? psi[2 * Np + i] : 0;
    for (int side = 0; side < 2; side++) {
        const int64_t j0 = (int64_t)nquad * (side == 0 ? idx_prev[i] : (int32_t)i);
        for (int q = 0; q < nquad; q++) {
            const int64_t j = j0 + q;
            const double rx = fold_sep(x - pts[j], box.L[0], box.Lh[0]);
            const double ry = fold_sep(y - pts[Nq + j], box.L[1], box.Lh[1]);
            const double rz = fold_sep(z - pts[2 * Nq + j], box.L[2], box.Lh[2]);
            const double r2 = dist2_nocontract(rx, ry, rz);
            const double r = sqrt(r2);
            const double rinv = 1.0 / r;
            double wa = 1.0, wb = 0.0;
            if (!full) {
                const double ar = alpha * r;
                wa = erf(ar);
                wb = TWO_OVER_SQRT_PI * ar * exp(-(ar * ar));
            }
            const double qx = chg[j], qy = chg[Nq + j], qz = chg[2 * Nq + j];
            if (do_vel) {
                const double f = (wa - wb) * (rinv * rinv * rinv);
                vx -= pref * (f * (qy * rz - qz * ry));
                vy -= pref * (f * (qz * rx - qx * rz));
                vz -= pref * (f * (qx * ry - qy * rx));
            }
            if (do_psi) {
                const double f = wa * rinv;
                px -= pref * (f * qx);
                py -= pref * (f * qy);
                pz -= pref * (f * qz);
            }
        }
    }
    if (do_vel) {
        v[i] = vx;
        v[Np + i] = vy;
        v[2 * Np + i] = vz;
    }
    if (do_psi) {
        psi[i] = px;
        psi[Np + i] = py;
        psi[2 * Np + i] = pz;
    }
}

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename CT, bool VEL, bool PSI>
void launch_pair(const PairArgs<CT> &a, bool stats, bool tiled, const int32_t *tstart, hipStream_t s) {
    if (tiled) {
        // one wavefront per tile of PT_T³ cells, padded for the XCD remap
        const int3 nt = make_int3((a.Nc.x + PT_T - 1) / PT_T, (a.Nc.y + PT_T - 1) / PT_T, (a.Nc.z + PT_T - 1) / PT_T);
        const int64_t ntiles = (int64_t)nt.x * nt.y * nt.z;
        const unsigned nb = (unsigned)(8 * ((ntiles + 7) / 8));
        if (stats)
            pair_tile_kernel<CT, VEL, PSI, true><<<nb, 64, 0, s>>>(a, tstart, nt);
        else
            pair_tile_kernel<CT, VEL, PSI, false><<<nb, 64, 0, s>>>(a, tstart, nt);
        return;
    }
    const unsigned nb = (unsigned)(8 * ((a.Np + 7) / 8));  // one wavefront per target, padded for the XCD remap
    if (stats)
        pair_wave_kernel<CT, VEL, PSI, true><<<nb, 64, 0, s>>>(a);
    else
        pair_wave_kernel<CT, VEL, PSI, false><<<nb, 64, 0, s>>>(a);
}

template <typename CT>
int run_pairs(vpbs_ctx *c, bool vel, bool psi, bool stats, bool include_local) {
    const int64_t Np = c->Np, Nq = c->Nq;
    hipStream_t s = c->s_sr;
    const int3 Nc = make_int3(c->Nc[0], c->Nc[1], c->Nc[2]);
    stage_begin(c, ST_CELL, s);
    cell_keys<<<nblocks(Nq, 256), 256, 0, s>>>(Nq, c->points, c->box, Nc, c->src_key, nullptr);
    VPBS_TRY(counting_sort(c, c->sort_sr, c->src_key, Nq, c->ncells, c->src_start, c->src_perm, s));
    gather_sources<CT><<<nblocks(Nq, 256), 256, 0, s>>>(Nq, c->src_perm, c->points, c->charges, c->src_pos,
                                                          (CT *)c->src_chg, c->src_idx);
    VPBS_TRY(build_target_cells(c, s));
    stage_end(c, ST_CELL, s);
    c->cells_version = c->points_version;
    if (c->lc_active && c->lc_ev_cells) VPBS_HIP(c, hipEventRecord(c->lc_ev_cells, s));  // learned check

    PairArgs<CT> a;
    a.Np = Np;
    a.Nq = Nq;
    a.tperm = c->tgt_perm;
    a.tcell = (const int4 *)c->tgt_cell;
    a.nodes = c->nodes;
    a.cstart = c->src_start;
    a.spos = c->src_pos;
    a.schg = (const CT *)c->src_chg;
    a.sidx = c->src_idx;
    a.idx_prev = c->idx_prev;
    a.nquad = c->p.nquad;
    a.box = c->box;
    a.Nc = Nc;
    a.M = c->p.nsubdiv;
    for (int d = 0; d < 3; d++) {
        a.cinv[d] = c->Nc[d] / c->p.Ls[d];
        a.cinvf[d] = (float)a.cinv[d];
        a.cszf[d] = (float)(c->p.Ls[d] / c->Nc[d]);
    }
    a.rcut2 = c->p.rcut * c->p.rcut;  // params.jl:107
    a.rcut2f = (float)a.rcut2;
    a.alpha = (CT)c->p.alpha;
    a.pref = c->p.Gamma / (4 * M_PI);
    a.include_local = include_local ? 1 : 0;
    a.v = c->v_sr;
    a.psi = c->psi_sr;
    a.counts = c->stat_counts;
    a.hashes = c->stat_hashes;
    a.cands = c->stat_cands;
    // the tiled kernel needs every cell of a tile's union exactly once: PT_T + 2M cells per direction
    const int M = a.M;
    const bool tiled = c->pair_tiled && Nc.x >= PT_T + 2 * M && Nc.y >= PT_T + 2 * M && Nc.z >= PT_T + 2 * M;
    stage_begin(c, ST_PAIR, s);
    if (vel && psi)
        launch_pair<CT, true, true>(a, stats, tiled, c->tgt_start, s);
    else if (vel)
        launch_pair<CT, true, false>(a, stats, tiled, c->tgt_start, s);
    else
        launch_pair<CT, false, true>(a, stats, tiled, c->tgt_start, s);
    stage_end(c, ST_PAIR, s);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

}  // namespace

// Nodes by cell of the short-range cell list. The order inside a cell is that of the atomic
// ranks: the pair kernel's sums do not depend on it (it only sets the processing order), and
// the feature kernel breaks distance ties by node index.
int build_target_cells(vpbs_ctx *c, hipStream_t s) {
    const int64_t Np = c->Np;
    const int3 Nc = make_int3(c->Nc[0], c->Nc[1], c->Nc[2]);
    cell_keys<<<nblocks(Np, 256), 256, 0, s>>>(Np, c->nodes, c->box, Nc, c->tgt_key, (int4 *)c->tgt_cell);
    VPBS_TRY(counting_sort(c, c->sort_tg, c->tgt_key, Np, c->ncells, c->tgt_start, c->tgt_perm, s, false));
    return VPBS_OK;
}

int shortrange_compute(vpbs_ctx *c, uint32_t flags, double integration_cutoff) {
    const int64_t Np = c->Np;
    const bool vel = flags & VPBS_VELOCITY, psi = flags & VPBS_STREAMFUNCTION;
    const bool lia_only = flags & VPBS_LIA_ONLY, lia_off = flags & VPBS_LIA_OFF;
    const bool stats = flags & VPBS_NEIGHBOUR_STATS;
    hipStream_t s = c->s_sr;
    const double pref = c->p.Gamma / (4 * M_PI);

    stage_begin(c, ST_LOCAL, s);
    if (!lia_off) {
        double cut_v, cut_psi;
        if (std::isnan(integration_cutoff)) {
            cut_psi = exp(c->p.Delta - 1) * c->p.a / 2;  // local_integrals.jl:118-119
            cut_v = cut_psi * exp(1.0);
        } else {
            cut_v = cut_psi = integration_cutoff;
        }
        local_term_kernel<<<nblocks(Np, 256), 256, 0, s>>>(Np, c->sp, c->spp, c->dlm, c->dlp, pref, cut_v, cut_psi,
                                                             vel, psi, c->v_sr, c->psi_sr);
    } else {
        if (vel) zero_f64<<<nblocks(3 * Np, 256), 256, 0, s>>>(c->v_sr, 3 * Np);
        if (psi) zero_f64<<<nblocks(3 * Np, 256), 256, 0, s>>>(c->psi_sr, 3 * Np);
    }
    stage_end(c, ST_LOCAL, s);
    VPBS_HIP(c, hipGetLastError());
    if (lia_only) return VPBS_OK;

    const bool fp32 = c->p.precision == VPBS_F32;
    // FP32 path always excludes the local segments (see header comment).
    const bool include_local = !fp32 && c->p.avoid_explicit_erf;
    if (fp32)
        VPBS_TRY(run_pairs<float>(c, vel, psi, stats, include_local));
    else
        VPBS_TRY(run_pairs<double>(c, vel, psi, stats, include_local));
    c->have_stats = stats;

    stage_begin(c, ST_SELF, s);
    self_kernel<<<nblocks(Np, 256), 256, 0, s>>>(Np, c->Nq, c->nodes, c->idx_prev, c->points, c->charges,
                                                 c->p.nquad, c->box, c->p.alpha, pref, include_local ? 1 : 0, vel,
                                                 psi, vel ? c->v_sr : nullptr, psi ? c->psi_sr : nullptr);
    stage_end(c, ST_SELF, s);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

}  // namespace vpbs
