// Learned correction of a cheap Ewald splitting (no counterpart in the reference).
//
// A context created with a cheap splitting (small β: short cut-off rcut_c, Ewald
// parameter α_c, coarse grid) is paired with a converged splitting on the same grid
// (α_f, rcut_f; β = 4 gives ~6e-9). The difference of the two velocities,
//
//   Δv = [pairs − local terms](α_f, rcut_f) − [pairs − local terms](α_c, rcut_c)
//        + Σ_k v̂(k) (e^{−k²/4α_f²} − e^{−k²/4α_c²}) e^{ik·x},
//
// is what the network predicts from local geometry. Both parts are computed exactly here
// (FP64 pair sums, the kept Fourier field v̂ with the difference filter) to make training
// targets and to check the network at run time:
//
//   strand_records    per node in cell-sorted order (the node cell list of the pair kernel):
//                     position, unit tangent and the offsets to the two neighbours along its
//                     filament (read as a candidate strand point), and the node's own set-up
//                     (local frame, filament, cell) read by the wavefront that builds its row
//   feature_kernel    one wavefront per node: local frame (t, n, b) from its set-up record; the K
//                     nearest strands within the feature radius, one point per strand (the
//                     node where the distance along that strand has a local minimum, moved
//                     to the foot of the perpendicular on its tangent line), with their
//                     tangents; all in the node's frame, lengths in units of 1/α_c
//   check_sr_kernel   one block per node: the short-range part of Δv over the sources within
//                     rcut_f (cells of the cheap cell list), FP64; the node's own two
//                     segments contribute −LR_f + LR_c (their long-range parts)
//   correction_apply  Δv_pred = (Γα_c/4π)·(y₀ t + y₁ n + y₂ b) from the network output y
//   check_reduce      sampled errors of the corrected and the cheap velocity: the sample's
//                     squared errors scaled to the whole field (×Np/ns) over Σ|v|² of every
//                     node (vnorm_partial), an estimate of the whole field's relative L2 error
//
// The correction is checked on a stratified random sample of nodes every call: when the
// relative L2 error of the corrected velocity against the converged one exceeds the
// tolerance, the correction is rejected and Δv is evaluated exactly on every node instead.
#include "vpbs_internal.hpp"
#include "device_util.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace vpbs {
namespace {

constexpr int LC_K = 16;        // strands per node
constexpr int LC_NIN = 128;     // feature row (6 + 6·K used)
constexpr int LC_QMAX = 128;    // queued strand points per node (reduced to the K nearest when it could overflow)
constexpr int LC_MAXROWS = 81;  // (2·4 + 1)² cell rows of the feature walk
constexpr double TWO_OVER_SQRT_PI = 1.1283791670955126;

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct Frame {
    double t[3], n[3], b[3], kappa;  // unit tangent, normal, binormal; curvature |κ|
};

// Local frame of a node from s′ and s″ (any parametrisation): t = s′/|s′|,
// κ = (s″ − (s″·t)t)/|s′|², n = κ/|κ|, b = t × n. Where |κ|/α is below 1e-9 the normal is
// t × e / |t × e| with e the axis least aligned with t.
__device__ inline Frame node_frame(const double *sp, const double *spp, int64_t Np, int64_t i, double alpha) {
    Frame f;
    const double s1[3] = {sp[i], sp[Np + i], sp[2 * Np + i]};
    const double s2[3] = {spp[i], spp[Np + i], spp[2 * Np + i]};
    const double m2 = s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2];
    const double mi = 1.0 / sqrt(m2);
    for (int d = 0; d < 3; d++) f.t[d] = s1[d] * mi;
    const double st = s2[0] * f.t[0] + s2[1] * f.t[1] + s2[2] * f.t[2];
    double k[3];
    for (int d = 0; d < 3; d++) k[d] = (s2[d] - st * f.t[d]) / m2;
    const double kn = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    f.kappa = kn;
    if (kn > 1e-9 * alpha) {
        for (int d = 0; d < 3; d++) f.n[d] = k[d] / kn;
    } else {
        const double a0 = fabs(f.t[0]), a1 = fabs(f.t[1]), a2 = fabs(f.t[2]);
        const int e = (a0 <= a1 && a0 <= a2) ? 0 : (a1 <= a2 ? 1 : 2);
        double u[3] = {0, 0, 0};
        u[e] = 1.0;
        double c[3] = {f.t[1] * u[2] - f.t[2] * u[1], f.t[2] * u[0] - f.t[0] * u[2], f.t[0] * u[1] - f.t[1] * u[0]};
        const double cn = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (int d = 0; d < 3; d++) f.n[d] = c[d] / cn;
    }
    f.b[0] = f.t[1] * f.n[2] - f.t[2] * f.n[1];
    f.b[1] = f.t[2] * f.n[0] - f.t[0] * f.n[2];
    f.b[2] = f.t[0] * f.n[1] - f.t[1] * f.n[0];
    return f;
}

// A node's own set-up for the feature row (one 128-byte record per slot, read with uniform
// loads by the node's wavefront): local frame t, n, b and |κ| (node_frame), the filament's
// first node and length, the node's place on it and its cell.
struct __align__(16) NodeSet {
    double f[10];  // t, n, b, |κ|
    int64_t oi;    // first node of the filament
    int32_t i, Ni, pi, pad;
    int4 cell;     // (ix, iy, iz, key) of the node cell list
};

// Per node, in cell-sorted order (slot k of the node cell list, node i = perm[k]): its
// position (FP64) and a record of float4 × 3: (t, i) with t the unit tangent s′/|s′| and the
// node index as bits, then w⁻ = s_{i−1} − s_i and w⁺ = s_{i+1} − s_i (minimum image, along
// i's filament). With r = s_i − x for a target x, |x − s_{i∓1}|² − |x − s_i|² = 2 r·w∓ + |w∓|²,
// so the local-minimum test of the features needs no neighbour loads. Also the node's set-up
// (NodeSet) and its frame in FP32 for the correction ([Np][9], original order).
__global__ void strand_records(int64_t Np, const int32_t *__restrict__ perm, const double *__restrict__ nodes,
                               const double *__restrict__ sp, const double *__restrict__ spp,
                               const int32_t *__restrict__ fid, const int64_t *__restrict__ foff,
                               const int32_t *__restrict__ fn, const int4 *__restrict__ tcell, Box box, double alpha,
                               double *__restrict__ npos, float4 *__restrict__ rec, NodeSet *__restrict__ nset,
                               float *__restrict__ frame) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Np) return;
    const int32_t i = perm[k];
    const double x = nodes[i], y = nodes[Np + i], z = nodes[2 * Np + i];
    npos[k] = x;
    npos[Np + k] = y;
    npos[2 * Np + k] = z;
    const int f = fid[i];
    const int64_t o = foff[f];
    const int n = fn[f], p = (int)(i - o);
    const int64_t im = o + (p + n - 1) % n, ip = o + (p + 1) % n;
    const float wmx = (float)fold_sep(nodes[im] - x, box.L[0], box.Lh[0]);
    const float wmy = (float)fold_sep(nodes[Np + im] - y, box.L[1], box.Lh[1]);
    const float wmz = (float)fold_sep(nodes[2 * Np + im] - z, box.L[2], box.Lh[2]);
    const float wpx = (float)fold_sep(nodes[ip] - x, box.L[0], box.Lh[0]);
    const float wpy = (float)fold_sep(nodes[Np + ip] - y, box.L[1], box.Lh[1]);
    const float wpz = (float)fold_sep(nodes[2 * Np + ip] - z, box.L[2], box.Lh[2]);
    const double s1x = sp[i], s1y = sp[Np + i], s1z = sp[2 * Np + i];
    const double nrm = sqrt(dist2_nocontract(s1x, s1y, s1z));
    rec[3 * k] = make_float4((float)(s1x / nrm), (float)(s1y / nrm), (float)(s1z / nrm), __int_as_float(i));
    rec[3 * k + 1] = make_float4(wmx, wmy, wmz, wpx);
    rec[3 * k + 2] = make_float4(wpy, wpz, 0.f, 0.f);
    const Frame F = node_frame(sp, spp, Np, i, alpha);
    NodeSet ns;
    for (int d = 0; d < 3; d++) {
        ns.f[d] = F.t[d];
        ns.f[3 + d] = F.n[d];
        ns.f[6 + d] = F.b[d];
        frame[(int64_t)i * 9 + d] = (float)F.t[d];
        frame[(int64_t)i * 9 + 3 + d] = (float)F.n[d];
        frame[(int64_t)i * 9 + 6 + d] = (float)F.b[d];
    }
    ns.f[9] = F.kappa;
    ns.oi = o;
    ns.i = i;
    ns.Ni = n;
    ns.pi = p;
    ns.pad = 0;
    ns.cell = tcell[i];
    nset[k] = ns;
}

struct FeatArgs {
    int64_t Np;
    const NodeSet *nset;    // cell-sorted node set-up (strand_records)
    const int32_t *tstart;  // node cell starts
    const double *npos;     // cell-sorted node positions [3][Np]
    const float4 *rec;      // cell-sorted strand records [Np][3] (strand_records)
    const double *nodes;
    Box box;
    int3 Nc;
    int Mf;
    int w2inv;    // ⌈65536 / (2·Mf + 1)⌉: row r of the walk is in z-layer (r · w2inv) >> 16
    double cinv[3];  // Nc / L: cells per unit length
    float cinvf[3], cszf[3];  // the same in FP32, and L / Nc (the rows' cut to the feature sphere)
    float rf2f;
    double rf2;   // feature radius²
    double alpha; // α_c (lengths in units of 1/α_c)
    float *feat;  // [Np][LC_NIN]
};

// One wavefront per node (cell-sorted order, XCD-aware block order).
__global__ __launch_bounds__(64) void feature_kernel(FeatArgs a) {
    // per row: (candidates before it, run 0 start, run 0 length, run 1 start - run 0 length)
    __shared__ int4 r_run[LC_MAXROWS];
    __shared__ float q_key[LC_QMAX];
    __shared__ int q_j[LC_QMAX];
    __shared__ float q_f[6][LC_QMAX];
    __shared__ float row[LC_NIN];
    const int lane = threadIdx.x;
    const int64_t Np = a.Np;
    const int64_t per = ((int64_t)gridDim.x + 7) / 8;
    const int64_t k = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (k >= Np) return;
    // the node's set-up: uniform loads, so the stretch and the cell rows below load at once
    const NodeSet *__restrict__ ns = a.nset + k;
    const int32_t i = ns->i;
    const double x[3] = {a.npos[k], a.npos[Np + k], a.npos[2 * Np + k]};
    const double Ft[3] = {ns->f[0], ns->f[1], ns->f[2]};
    const double Fn[3] = {ns->f[3], ns->f[4], ns->f[5]};
    const double Fb[3] = {ns->f[6], ns->f[7], ns->f[8]};
    const double kappa = ns->f[9];
    const int64_t oi = ns->oi;
    const int Ni = ns->Ni;
    const int pi = ns->pi;

    // the node's own stretch: consecutive neighbours along its filament within the radius
    int dplus, dminus;
    {
        const int step = lane < 32 ? lane + 1 : -(lane - 31);
        int p = pi + step;
        if (Ni > 32) {  // one wrap at most
            p += (p < 0) ? Ni : 0;
            p -= (p >= Ni) ? Ni : 0;
        } else {
            p = (p % Ni + Ni) % Ni;
        }
        const int64_t j = oi + p;
        const double r0 = fold_sep(a.nodes[j] - x[0], a.box.L[0], a.box.Lh[0]);
        const double r1 = fold_sep(a.nodes[Np + j] - x[1], a.box.L[1], a.box.Lh[1]);
        const double r2 = fold_sep(a.nodes[2 * Np + j] - x[2], a.box.L[2], a.box.Lh[2]);
        const bool beyond = dist2_nocontract(r0, r1, r2) > a.rf2;
        const unsigned long long m = __ballot(beyond);
        const unsigned mf = (unsigned)(m & 0xffffffffull), mb = (unsigned)(m >> 32);
        dplus = mf ? __builtin_ctz(mf) : 32;
        dminus = mb ? __builtin_ctz(mb) : 32;
    }

    // cell rows of the walk: lanes take rows lane and lane + 64 (≤ 2 runs each with the
    // periodic wrap along x); a wave scan gives each row's first candidate. Each row keeps
    // only the cells that can hold a point within the radius: rows whose (y, z) gap to the
    // node exceeds it are empty, the others are cut to the x-cells within
    // √(r_f² − gap²) of the node (in FP32, widened by 1e-4 cells: a conservative cut; the
    // distance test below decides).
    const int4 tc = ns->cell;
    const int M = a.Mf, W2 = 2 * M + 1, nrows = W2 * W2;
    // The minimum-image fold changes no separation to a candidate when the node's neighbour
    // cells do not wrap and every direction has at least 2M + 2 cells: then |x − s| < (M + 1)
    // cells ≤ L/2 (up to rounding at exactly L/2, beyond the radius either way). Such nodes
    // (most of them) skip the fold; the own-stretch test above keeps it (neighbours along a
    // filament may sit across the periodic boundary).
    const bool wraps = tc.x < M || tc.y < M || tc.z < M || tc.x + M >= a.Nc.x || tc.y + M >= a.Nc.y ||
                       tc.z + M >= a.Nc.z || min(min(a.Nc.x, a.Nc.y), a.Nc.z) < 2 * M + 2;
    // the node's place in its cell (in cells; within a few ulps of the cell index's own
    // (x / L)·Nc, and within FP32 rounding after the cast: far inside the 1e-4 widening)
    const float u[3] = {(float)(x[0] * a.cinv[0] - tc.x), (float)(x[1] * a.cinv[1] - tc.y),
                        (float)(x[2] * a.cinv[2] - tc.z)};
    int len[2] = {0, 0};
    int4 run[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int r = lane + 64 * h;
        if (r < nrows) {
            const int qz = (r * a.w2inv) >> 16, dz = qz - M, dy = (r - qz * W2) - M;
            int zz = tc.z + dz, yy = tc.y + dy;
            zz += (zz < 0) ? a.Nc.z : 0;
            zz -= (zz >= a.Nc.z) ? a.Nc.z : 0;
            yy += (yy < 0) ? a.Nc.y : 0;
            yy -= (yy >= a.Nc.y) ? a.Nc.y : 0;
            const int64_t rb = ((int64_t)zz * a.Nc.y + yy) * a.Nc.x;
            const float gy = fmaxf((dy > 0 ? dy - u[1] : (dy < 0 ? u[1] - (dy + 1) : 0.f)) - 1e-4f, 0.f) * a.cszf[1];
            const float gz = fmaxf((dz > 0 ? dz - u[2] : (dz < 0 ? u[2] - (dz + 1) : 0.f)) - 1e-4f, 0.f) * a.cszf[2];
            const float rem = a.rf2f - (gy * gy + gz * gz);
            int lo = 1, hi = 0;  // empty
            if (rem >= 0.f) {
                const float ext = sqrtf(rem) * a.cinvf[0] + 1e-4f;
                lo = tc.x + max((int)floorf(u[0] - ext), -M);
                hi = tc.x + min((int)floorf(u[0] + ext), M);
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
                s0 = a.tstart[rb + ca0];
                n0 = a.tstart[rb + cb0 + 1] - s0;
            }
            if (cb1 >= ca1) {
                s1 = a.tstart[rb + ca1];
                n1 = a.tstart[rb + cb1 + 1] - s1;
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
    const int ncand = nrows > 64 ? __builtin_amdgcn_readlane(incl1, nrows - 65)
                                 : __builtin_amdgcn_readlane(incl0, nrows - 1);
    int *const s_mark = reinterpret_cast<int *>(row);  // row marks of the batch (row[] is zeroed after the walk)
    __syncthreads();

    // Candidate b + lane: its row from markers (each non-empty row that meets the batch marks
    // its first candidate in it; a max-scan carries the marks forward), then its slot, record
    // and position, loaded one batch ahead.
    auto fetch = [&](int b, int &fs) {
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
    };
    int s = 0;
    if (ncand > 0) fetch(0, s);
    float4 ra = a.rec[3 * s], rb = a.rec[3 * s + 1], rc = a.rec[3 * s + 2];
    double px0 = a.npos[s], px1 = a.npos[Np + s], px2 = a.npos[2 * Np + s];

    int nq = 0;  // queued strand points (uniform)
    for (int base = 0; base < ncand; base += 64) {
        int s_n;
        fetch(base + 64, s_n);
        const float4 ra_n = a.rec[3 * s_n], rb_n = a.rec[3 * s_n + 1], rc_n = a.rec[3 * s_n + 2];
        const double px0_n = a.npos[s_n], px1_n = a.npos[Np + s_n], px2_n = a.npos[2 * Np + s_n];
        const int cnd = base + lane;
        bool acc = false;
        float key = 0.f, fv[6] = {0, 0, 0, 0, 0, 0};
        const int32_t j = __float_as_int(ra.w);
        {
            double r0 = px0 - x[0], r1 = px1 - x[1], r2v = px2 - x[2];
            if (wraps) {  // wave-uniform
                r0 = fold_sep(r0, a.box.L[0], a.box.Lh[0]);
                r1 = fold_sep(r1, a.box.L[1], a.box.Lh[1]);
                r2v = fold_sep(r2v, a.box.L[2], a.box.Lh[2]);
            }
            const double d2 = dist2_nocontract(r0, r1, r2v);
            bool ok = cnd < ncand && j != i && d2 <= a.rf2;
            if (ok && (unsigned)(j - oi) < (unsigned)Ni) {  // i's own filament
                const int fo = (j - i) + ((j < i) ? Ni : 0);  // steps forward from i to j
                if (fo <= dplus || Ni - fo <= dminus) ok = false;
            }
            const float rx = (float)r0, ry = (float)r1, rz = (float)r2v;
            if (ok) {
                // local minimum of the distance along j's filament: d²(j−1) ≥ d²(j) < d²(j+1)
#pragma clang fp contract(off)
                const float gm = 2.f * ((rx * rb.x + ry * rb.y) + rz * rb.z) + ((rb.x * rb.x + rb.y * rb.y) + rb.z * rb.z);
                const float gp = 2.f * ((rx * rb.w + ry * rc.x) + rz * rc.y) + ((rb.w * rb.w + rc.x * rc.x) + rc.y * rc.y);
                ok = gm >= 0.f && gp > 0.f;
            }
            if (ok) {
                // foot of the perpendicular on the strand's tangent line; projected onto the
                // frame only for the K strands kept
#pragma clang fp contract(off)
                const float rt = (rx * ra.x + ry * ra.y) + rz * ra.z;
                const float qx = rx - rt * ra.x, qy = ry - rt * ra.y, qz = rz - rt * ra.z;
                fv[0] = qx, fv[1] = qy, fv[2] = qz;
                fv[3] = ra.x, fv[4] = ra.y, fv[5] = ra.z;
                key = (qx * qx + qy * qy) + qz * qz;
                acc = true;
            }
        }
        ra = ra_n, rb = rb_n, rc = rc_n;
        px0 = px0_n, px1 = px1_n, px2 = px2_n;
        const unsigned long long mask = __ballot(acc);
        if (acc) {
            const int pos = nq + __popcll(mask & ((1ull << lane) - 1ull));
            {  // pos < LC_QMAX: the queue holds at most LC_QMAX − 64 entries before a batch
                q_key[pos] = key;
                q_j[pos] = j;
#pragma unroll
                for (int e = 0; e < 6; e++) q_f[e][pos] = fv[e];
            }
        }
        nq += __popcll(mask);
        if (nq > LC_QMAX - 64) {
            // the next batch could overflow the queue: keep its K nearest entries (by
            // perpendicular distance, ties by node index), in that order. An entry with K
            // others nearer can never be among the K nearest of the node, so the kept set
            // is exact whatever the density.
            wave_lds_order();
            int rk[2] = {LC_K, LC_K};
            float ck[2] = {0.f, 0.f}, cf[2][6];
            int cj[2] = {0, 0};
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int e = lane + 64 * h;
                if (e < nq) {
                    ck[h] = q_key[e];
                    cj[h] = q_j[e];
#pragma unroll
                    for (int f = 0; f < 6; f++) cf[h][f] = q_f[f][e];
                    int r = 0;
                    for (int o = 0; o < nq; o++) {
                        const float ko = q_key[o];
                        r += (ko < ck[h]) || (ko == ck[h] && q_j[o] < cj[h]);
                    }
                    rk[h] = r;
                }
            }
            wave_lds_order();
#pragma unroll
            for (int h = 0; h < 2; h++)
                if (rk[h] < LC_K) {
                    q_key[rk[h]] = ck[h];
                    q_j[rk[h]] = cj[h];
#pragma unroll
                    for (int f = 0; f < 6; f++) q_f[f][rk[h]] = cf[h][f];
                }
            nq = LC_K;
            wave_lds_order();
        }
    }
    wave_lds_order();  // the row marks (ints in row's memory) are read before row is zeroed
    for (int e = lane; e < LC_NIN; e += 64) row[e] = 0.f;
    __syncthreads();
    nq = nq < LC_QMAX ? nq : LC_QMAX;
    // the K nearest strands (by perpendicular distance, ties by node index) in order
    for (int e = lane; e < nq; e += 64) {
        const float ke = q_key[e];
        const int je = q_j[e];
        int rank = 0;
        for (int o = 0; o < nq; o++) {
            const float ko = q_key[o];
            rank += (ko < ke) || (ko == ke && q_j[o] < je);
        }
        if (rank < LC_K) {
#pragma clang fp contract(off)
            const float qx = q_f[0][e], qy = q_f[1][e], qz = q_f[2][e];
            const float tx = q_f[3][e], ty = q_f[4][e], tz = q_f[5][e];
            float *rr = row + 6 + 6 * rank;
            rr[0] = (float)(a.alpha * ((qx * Ft[0] + qy * Ft[1]) + qz * Ft[2]));
            rr[1] = (float)(a.alpha * ((qx * Fn[0] + qy * Fn[1]) + qz * Fn[2]));
            rr[2] = (float)(a.alpha * ((qx * Fb[0] + qy * Fb[1]) + qz * Fb[2]));
            rr[3] = (float)((tx * Ft[0] + ty * Ft[1]) + tz * Ft[2]);
            rr[4] = (float)((tx * Fn[0] + ty * Fn[1]) + tz * Fn[2]);
            rr[5] = (float)((tx * Fb[0] + ty * Fb[1]) + tz * Fb[2]);
        }
    }
    if (lane == 0) {
        row[1] = (float)(kappa / a.alpha);    // curvature vector in the frame: (0, |κ|/α, 0)
        row[3] = 1.f;                         // tangent in the frame: (1, 0, 0)
    }
    __syncthreads();
    float4 *dst = (float4 *)(a.feat + (int64_t)i * LC_NIN);
    if (lane < LC_NIN / 4) dst[lane] = make_float4(row[4 * lane], row[4 * lane + 1], row[4 * lane + 2], row[4 * lane + 3]);
}

struct CheckArgs {
    int64_t Np, Nq, n;
    const int32_t *list;   // nodes to evaluate (null: all, in order)
    const double *nodes;
    const int32_t *idx_prev;
    const int4 *tcell;
    const int32_t *cstart;  // source cell starts
    const double *spos;     // cell-sorted source positions [3][Nq]
    const int32_t *sidx;    // original index of each sorted source
    const double *pts, *chg;  // original order [3][Nq]
    int nquad;
    Box box;
    int3 Nc, Mw;
    double cinv[3];  // Nc / L: cells per unit length
    double csz[3];   // L / Nc: cell sizes
    double rc2, rf2, alpha_c, alpha_f, pref;
    double *out;  // [3][n]
};

__device__ inline double srw(double alpha, double r) {  // short-range velocity weight
    const double ar = alpha * r;
    return erfc(ar) + TWO_OVER_SQRT_PI * ar * exp(-ar * ar);
}
__device__ inline double lrw(double alpha, double r) {  // long-range velocity weight
    const double ar = alpha * r;
    return erf(ar) - TWO_OVER_SQRT_PI * ar * exp(-ar * ar);
}

constexpr int CHK_T = 256;
constexpr int CHK_MAXROWS = 33 * 33;  // 2·Mw + 1 ≤ 33

// One block per node: short-range part of Δv (converged − cheap), FP64.
__global__ __launch_bounds__(CHK_T) void check_sr_kernel(CheckArgs a) {
    __shared__ int r_pre[CHK_MAXROWS + 1];
    __shared__ int r_s0[CHK_MAXROWS], r_n0[CHK_MAXROWS], r_s1[CHK_MAXROWS];
    __shared__ double red[3][CHK_T];
    const int tid = threadIdx.x;
    const int64_t k = blockIdx.x;
    if (k >= a.n) return;
    const int64_t Np = a.Np, Nq = a.Nq;
    const int32_t i = a.list ? a.list[k] : (int32_t)k;
    const double x[3] = {a.nodes[i], a.nodes[Np + i], a.nodes[2 * Np + i]};
    const int4 tc = a.tcell[i];
    const int WY = 2 * a.Mw.y + 1, WZ = 2 * a.Mw.z + 1, nrows = WY * WZ;
    // rows cut to the cells that the rcut_f sphere reaches (conservative by 1e-9 cells; the
    // r² test decides), as in the pair kernel
    const double u[3] = {x[0] * a.cinv[0] - tc.x, x[1] * a.cinv[1] - tc.y, x[2] * a.cinv[2] - tc.z};
    for (int r = tid; r < nrows; r += CHK_T) {
        const int dz = r / WY - a.Mw.z, dy = r % WY - a.Mw.y;
        int zz = tc.z + dz, yy = tc.y + dy;
        zz += (zz < 0) ? a.Nc.z : 0;
        zz -= (zz >= a.Nc.z) ? a.Nc.z : 0;
        yy += (yy < 0) ? a.Nc.y : 0;
        yy -= (yy >= a.Nc.y) ? a.Nc.y : 0;
        const int64_t rb = ((int64_t)zz * a.Nc.y + yy) * a.Nc.x;
        const double gy = fmax((dy > 0 ? dy - u[1] : (dy < 0 ? u[1] - (dy + 1) : 0.0)) - 1e-9, 0.0) * a.csz[1];
        const double gz = fmax((dz > 0 ? dz - u[2] : (dz < 0 ? u[2] - (dz + 1) : 0.0)) - 1e-9, 0.0) * a.csz[2];
        const double rem = a.rf2 - (gy * gy + gz * gz);
        int lo = 1, hi = 0;  // empty
        if (rem >= 0.0) {
            const double ext = sqrt(rem) * a.cinv[0] + 1e-9;
            lo = tc.x + max((int)floor(u[0] - ext), -a.Mw.x);
            hi = tc.x + min((int)floor(u[0] + ext), a.Mw.x);
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
            s0 = a.cstart[rb + ca0];
            n0 = a.cstart[rb + cb0 + 1] - s0;
        }
        if (cb1 >= ca1) {
            s1 = a.cstart[rb + ca1];
            n1 = a.cstart[rb + cb1 + 1] - s1;
        }
        r_s0[r] = s0;
        r_n0[r] = n0;
        r_s1[r] = s1;
        r_pre[r + 1] = n0 + n1;
    }
    __syncthreads();
    // prefix over the rows: each wavefront scans 64 rows at a time, carrying the total
    if (tid < 64) {
        int carry = 0;
        for (int r0 = 0; r0 < nrows; r0 += 64) {
            const int r = r0 + tid;
            const int v = r < nrows ? r_pre[r + 1] : 0;
            const int incl = wave_incl_scan(v) + carry;
            if (r < nrows) r_pre[r + 1] = incl;
            carry = __builtin_amdgcn_readlane(incl, 63);
        }
        if (tid == 0) r_pre[0] = 0;
    }
    __syncthreads();
    const int ncand = r_pre[nrows];
    const int jr0 = a.nquad * i, jl0 = a.nquad * a.idx_prev[i];
    double acc[3] = {0, 0, 0};
    for (int cnd = tid; cnd < ncand; cnd += CHK_T) {
        int lo = 0, hi = nrows - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (r_pre[mid] <= cnd) lo = mid; else hi = mid - 1;
        }
        const int o = cnd - r_pre[lo];
        const int s = o < r_n0[lo] ? r_s0[lo] + o : r_s1[lo] + (o - r_n0[lo]);
        const int32_t j = a.sidx[s];
        if ((unsigned)(j - jr0) < (unsigned)a.nquad || (unsigned)(j - jl0) < (unsigned)a.nquad) continue;
        const double rx = fold_sep(x[0] - a.spos[s], a.box.L[0], a.box.Lh[0]);
        const double ry = fold_sep(x[1] - a.spos[Nq + s], a.box.L[1], a.box.Lh[1]);
        const double rz = fold_sep(x[2] - a.spos[2 * Nq + s], a.box.L[2], a.box.Lh[2]);
        const double r2 = dist2_nocontract(rx, ry, rz);
        if (!(r2 <= a.rf2)) continue;
        const double r = sqrt(r2);
        double w = srw(a.alpha_f, r);
        if (r2 <= a.rc2) w -= srw(a.alpha_c, r);
        const double f = w / (r2 * r);
        const double qx = a.chg[j], qy = a.chg[Nq + j], qz = a.chg[2 * Nq + j];
        acc[0] += f * (qy * rz - qz * ry);
        acc[1] += f * (qz * rx - qx * rz);
        acc[2] += f * (qx * ry - qy * rx);
    }
    // the node's two segments: their long-range parts, −LR(α_f) + LR(α_c)
    if (tid < 2 * a.nquad) {
        const int64_t j = tid < a.nquad ? jr0 + tid : jl0 + (tid - a.nquad);
        const double rx = fold_sep(x[0] - a.pts[j], a.box.L[0], a.box.Lh[0]);
        const double ry = fold_sep(x[1] - a.pts[Nq + j], a.box.L[1], a.box.Lh[1]);
        const double rz = fold_sep(x[2] - a.pts[2 * Nq + j], a.box.L[2], a.box.Lh[2]);
        const double r2 = dist2_nocontract(rx, ry, rz);
        const double r = sqrt(r2);
        const double f = (lrw(a.alpha_c, r) - lrw(a.alpha_f, r)) / (r2 * r);
        const double qx = a.chg[j], qy = a.chg[Nq + j], qz = a.chg[2 * Nq + j];
        acc[0] += f * (qy * rz - qz * ry);
        acc[1] += f * (qz * rx - qx * rz);
        acc[2] += f * (qx * ry - qy * rx);
    }
    for (int d = 0; d < 3; d++) red[d][tid] = acc[d];
    __syncthreads();
    for (int w = CHK_T / 2; w > 0; w >>= 1) {
        if (tid < w)
            for (int d = 0; d < 3; d++) red[d][tid] += red[d][tid + w];
        __syncthreads();
    }
    if (tid < 3) a.out[tid * a.n + k] = a.pref * red[tid][0];
}

// Stratified random sample: node k·Np/ns + u·Np/ns, u from a hash of (seed, k).
__global__ void draw_sample(int64_t Np, int64_t ns, uint64_t seed, const double *__restrict__ nodes,
                            int32_t *__restrict__ sidx, double *__restrict__ spos) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const double u = (double)(mix64(seed ^ mix64((uint64_t)k)) >> 11) * 0x1.0p-53;
    int64_t i = (int64_t)(((double)k + u) * (double)Np / (double)ns);
    i = i < Np ? i : Np - 1;
    sidx[k] = (int32_t)i;
#pragma unroll
    for (int d = 0; d < 3; d++) spos[d * ns + k] = nodes[d * Np + i];
}

// Δv_pred from the network output y (frame components, units Γα_c/4π); v += Δv_pred.
__global__ void correction_apply(int64_t Np, const float *__restrict__ y, const float *__restrict__ frame, double U,
                                 double *__restrict__ dvp, double *__restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    const float *F = frame + i * 9;
    const double y0 = y[i * 3], y1 = y[i * 3 + 1], y2 = y[i * 3 + 2];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double dv = U * (y0 * F[d] + y1 * F[3 + d] + y2 * F[6 + d]);
        dvp[d * Np + i] = dv;
        v[d * Np + i] += dv;
    }
}

// Sums over the sample of |Δv_pred − Δv|², |Δv|² and |v_cheap + Δv|² (fixed-order tree).
// Per-block partial sums of |v|² over every node (grid-stride, at most CHK_NV blocks), added by
// check_reduce's fixed reduction tree (one partial per thread).
constexpr int CHK_NV = 1024;
static_assert(CHK_NV <= 1024, "one partial per thread of check_reduce");
__global__ __launch_bounds__(256) void vnorm_partial(int64_t Np, const double *__restrict__ v,
                                                     double *__restrict__ part) {
    __shared__ double red[256];
    double a = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < Np; i += (int64_t)gridDim.x * 256)
        a += v[i] * v[i] + v[Np + i] * v[Np + i] + v[2 * Np + i] * v[2 * Np + i];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out = (Σ_s |Δv_pred − Δv|², Σ_s |Δv|², (ns/Np)·Σ_all |v|²) over the sample s: the ratios of the
// first two to the third estimate the squared relative L2 errors of the corrected and the cheap
// velocity over the whole field. (Normalising by the sample's own Σ|v|² would not: a few fast
// nodes hold much of the field's Σ|v|², and a 1 % sample that misses them overstates the
// relative error two- to four-fold.)
__global__ __launch_bounds__(1024) void check_reduce(int64_t Np, int64_t ns, const int32_t *__restrict__ sidx,
                                                    const double *__restrict__ dsr, const double *__restrict__ dlr,
                                                    const double *__restrict__ dvp, const double *__restrict__ vpart,
                                                    int nvp, double *__restrict__ out) {
    __shared__ double red[3][1024];
    double a0 = 0, a1 = 0;
    for (int64_t k = threadIdx.x; k < ns; k += 1024) {
        const int64_t i = sidx[k];
        for (int d = 0; d < 3; d++) {
            const double dt = dsr[d * ns + k] + dlr[d * ns + k];
            const double dp = dvp ? dvp[d * Np + i] : 0.0;
            a0 += (dp - dt) * (dp - dt);
            a1 += dt * dt;
        }
    }
    red[0][threadIdx.x] = a0;
    red[1][threadIdx.x] = a1;
    red[2][threadIdx.x] = (int)threadIdx.x < nvp ? vpart[threadIdx.x] : 0.0;  // nvp ≤ 1024
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 2) out[threadIdx.x] = red[threadIdx.x][0];
    if (threadIdx.x == 2) out[2] = red[2][0] * ((double)ns / (double)Np);
}

// v ← v − Δv_pred + Δv (exact difference on every node).
__global__ void correction_exact(int64_t Np, const double *__restrict__ dvp, const double *__restrict__ dsr,
                                 const double *__restrict__ dlr, double *__restrict__ v) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 3 * Np) return;
    v[e] += dsr[e] + dlr[e] - (dvp ? dvp[e] : 0.0);
}

// Training target: Δv in the node's frame, in units of Γα_c/4π ([Np][3] FP32).
__global__ void target_kernel(int64_t Np, const double *__restrict__ sp, const double *__restrict__ spp,
                              double alpha, double U_inv, const double *__restrict__ dsr,
                              const double *__restrict__ dlr, float *__restrict__ T) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    const Frame F = node_frame(sp, spp, Np, i, alpha);
    double dv[3];
    for (int d = 0; d < 3; d++) dv[d] = dsr[d * Np + i] + dlr[d * Np + i];
    T[i * 3] = (float)(U_inv * (dv[0] * F.t[0] + dv[1] * F.t[1] + dv[2] * F.t[2]));
    T[i * 3 + 1] = (float)(U_inv * (dv[0] * F.n[0] + dv[1] * F.n[1] + dv[2] * F.n[2]));
    T[i * 3 + 2] = (float)(U_inv * (dv[0] * F.b[0] + dv[1] * F.b[1] + dv[2] * F.b[2]));
}

template <typename T>
int regrow(vpbs_ctx *c, T **p, int64_t n) {
    if (*p) VPBS_HIP(c, hipFree(*p));
    *p = nullptr;
    VPBS_HIP(c, hipMalloc((void **)p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    return VPBS_OK;
}

int lc_buffers(vpbs_ctx *c) {
    const int64_t Np = c->Np;
    if (Np > c->lc_cap) {
        VPBS_TRY(regrow(c, &c->lc_npos, 3 * Np));
        VPBS_TRY(regrow(c, &c->lc_rec, 12 * Np));
        VPBS_TRY(regrow(c, (NodeSet **)&c->lc_nset, Np));
        VPBS_TRY(regrow(c, &c->lc_feat, LC_NIN * Np));
        VPBS_TRY(regrow(c, &c->lc_frame, 9 * Np));
        VPBS_TRY(regrow(c, &c->lc_y, 3 * Np));
        VPBS_TRY(regrow(c, &c->lc_dvp, 3 * Np));
        VPBS_TRY(regrow(c, &c->lc_dsr, 3 * Np));
        VPBS_TRY(regrow(c, &c->lc_dlr, 3 * Np));
        c->lc_cap = Np;
    }
    if (!c->lc_stats) VPBS_TRY(regrow(c, &c->lc_stats, 4 + CHK_NV));
    if (!c->lc_ev) VPBS_HIP(c, hipEventCreateWithFlags(&c->lc_ev, hipEventDisableTiming));
    if (!c->lc_ev_cells) VPBS_HIP(c, hipEventCreateWithFlags(&c->lc_ev_cells, hipEventDisableTiming));
    return VPBS_OK;
}

int sample_buffers(vpbs_ctx *c, int64_t ns) {
    if (ns > c->lc_cap_s) {
        VPBS_TRY(regrow(c, &c->lc_sidx, ns));
        VPBS_TRY(regrow(c, &c->lc_spos, 3 * ns));
        c->lc_cap_s = ns;
    }
    return VPBS_OK;
}

// Features of the current point data (node cell list current; short-range stream).
int run_features(vpbs_ctx *c, hipStream_t s) {
    const int64_t Np = c->Np;
    VPBS_TRY(lc_buffers(c));
    strand_records<<<nblk(Np, 256), 256, 0, s>>>(Np, c->tgt_perm, c->nodes, c->sp, c->spp, c->fil_fid, c->fil_off,
                                                 c->fil_n, (const int4 *)c->tgt_cell, c->box, c->lc_alpha_c,
                                                 c->lc_npos, (float4 *)c->lc_rec, (NodeSet *)c->lc_nset,
                                                 c->lc_frame);
    FeatArgs a;
    a.Np = Np;
    a.nset = (const NodeSet *)c->lc_nset;
    a.tstart = c->tgt_start;
    a.npos = c->lc_npos;
    a.rec = (const float4 *)c->lc_rec;
    a.nodes = c->nodes;
    a.box = c->box;
    a.Nc = make_int3(c->Nc[0], c->Nc[1], c->Nc[2]);
    int Mf = 0;
    for (int d = 0; d < 3; d++) Mf = std::max(Mf, (int)std::ceil(c->lc_rfeat * c->Nc[d] / c->p.Ls[d] - 1e-12));
    a.Mf = Mf;
    a.w2inv = (65536 + 2 * Mf) / (2 * Mf + 1);
    for (int d = 0; d < 3; d++) {
        a.cinv[d] = c->Nc[d] / c->p.Ls[d];
        a.cinvf[d] = (float)a.cinv[d];
        a.cszf[d] = (float)(c->p.Ls[d] / c->Nc[d]);
    }
    a.rf2 = c->lc_rfeat * c->lc_rfeat;
    a.rf2f = (float)a.rf2;
    a.alpha = c->lc_alpha_c;
    a.feat = c->lc_feat;
    feature_kernel<<<(unsigned)(8 * ((Np + 7) / 8)), 64, 0, s>>>(a);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

// Short-range part of Δv at n nodes (list, or all nodes when list is null) into out [3][n].
int run_check_sr(vpbs_ctx *c, int64_t n, const int32_t *list, double *out, hipStream_t s) {
    if (n <= 0) return VPBS_OK;
    CheckArgs a;
    a.Np = c->Np;
    a.Nq = c->Nq;
    a.n = n;
    a.list = list;
    a.nodes = c->nodes;
    a.idx_prev = c->idx_prev;
    a.tcell = (const int4 *)c->tgt_cell;
    a.cstart = c->src_start;
    a.spos = c->src_pos;
    a.sidx = c->src_idx;
    a.pts = c->points;
    a.chg = c->charges;
    a.nquad = c->p.nquad;
    a.box = c->box;
    a.Nc = make_int3(c->Nc[0], c->Nc[1], c->Nc[2]);
    int mw[3];
    for (int d = 0; d < 3; d++) mw[d] = (int)std::ceil(c->lc_rcut_f * c->Nc[d] / c->p.Ls[d] - 1e-12);
    a.Mw = make_int3(mw[0], mw[1], mw[2]);
    for (int d = 0; d < 3; d++) {
        a.cinv[d] = c->Nc[d] / c->p.Ls[d];
        a.csz[d] = c->p.Ls[d] / c->Nc[d];
    }
    a.rc2 = c->p.rcut * c->p.rcut;
    a.rf2 = c->lc_rcut_f * c->lc_rcut_f;
    a.alpha_c = c->p.alpha;
    a.alpha_f = c->lc_alpha_f;
    a.pref = c->p.Gamma / (4 * M_PI);
    a.out = out;
    check_sr_kernel<<<(unsigned)n, CHK_T, 0, s>>>(a);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

// Long-range part of Δv at n positions pos [3][n] into out [3][n] (long-range stream).
int run_check_lr(vpbs_ctx *c, int64_t n, const double *pos, double *out) {
    const double cf = 1.0 / (4 * c->lc_alpha_f * c->lc_alpha_f), cc = 1.0 / (4 * c->p.alpha * c->p.alpha);
    VPBS_TRY(longrange_filtered_grid(c, cf, cc));
    return longrange_interp_at(c, n, pos, out);
}

int state_check(vpbs_ctx *c, bool need_uhat) {
    if (!c->lc_ready) return fail(c, VPBS_ERR_STATE, "learned correction not set up: call vpbs_learned_setup");
    if (!c->have_points || !c->have_filaments)
        return fail(c, VPBS_ERR_STATE, "the learned correction needs filaments on the device (vpbs_set_filaments)");
    if (c->cells_version != c->points_version)
        return fail(c, VPBS_ERR_STATE, "cell lists are not those of the current points: run vpbs_compute first");
    if (need_uhat && (c->uhat_version != c->points_version || c->uhat_quantity != VPBS_UHAT_VELOCITY))
        return fail(c, VPBS_ERR_STATE,
                    "no kept velocity field for the current points: run vpbs_compute with VPBS_VELOCITY | "
                    "VPBS_KEEP_UHAT first");
    return VPBS_OK;
}

}  // namespace

// Inside vpbs_compute with the learned correction active: features and network inference on
// the short-range stream (beside the long range).
int learned_after_shortrange(vpbs_ctx *c) {
    hipStream_t s = c->s_sr;
    stage_begin(c, ST_FEATURES, s);
    VPBS_TRY(run_features(c, s));
    stage_end(c, ST_FEATURES, s);
    stage_begin(c, ST_MLP, s);
    const int rc = mlp_infer(c->lc_net, c->Np, c->lc_feat, c->lc_y, s);
    if (rc != VPBS_OK) return fail(c, rc, "network inference failed: %s", mlp_error(c->lc_net));
    stage_end(c, ST_MLP, s);
    return VPBS_OK;
}

// Inside vpbs_compute (learned evaluation with a sample): the long-range part of the sampled
// check on the long-range stream, right after the long range (the kept field is ready there).
int learned_check_long(vpbs_ctx *c) {
    if (c->lc_ns <= 0) return VPBS_OK;
    stage_begin(c, ST_CHECK, c->s_lr);
    return run_check_lr(c, c->lc_ns, c->lc_spos, c->lc_dlr);
}

// ... and its short-range part, also on the long-range stream, once the short-range stream
// has built the cell lists (event recorded after the cell build).
int learned_check_short(vpbs_ctx *c) {
    if (c->lc_ns <= 0) return VPBS_OK;
    VPBS_HIP(c, hipStreamWaitEvent(c->s_lr, c->lc_ev_cells, 0));
    VPBS_TRY(run_check_sr(c, c->lc_ns, c->lc_sidx, c->lc_dsr, c->s_lr));
    stage_end(c, ST_CHECK, c->s_lr);
    return VPBS_OK;
}

// After the merge: the predicted correction added to the velocity.
int learned_after_merge(vpbs_ctx *c) {
    hipStream_t s = c->s_sr;
    const double U = c->p.Gamma * c->lc_alpha_c / (4 * M_PI);
    correction_apply<<<nblk(c->Np, 256), 256, 0, s>>>(c->Np, c->lc_y, c->lc_frame, U, c->lc_dvp, c->v_out);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

void learned_release(vpbs_ctx *c) {
    void *bufs[] = {c->lc_npos, c->lc_rec, c->lc_nset, c->lc_feat, c->lc_frame, c->lc_y, c->lc_dvp, c->lc_dsr, c->lc_dlr, c->lc_sidx,
                    c->lc_spos, c->lc_stats};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->lc_ev) (void)hipEventDestroy(c->lc_ev);
    if (c->lc_ev_cells) (void)hipEventDestroy(c->lc_ev_cells);
}

}  // namespace vpbs

using namespace vpbs;

extern "C" {

int vpbs_learned_setup(vpbs_ctx *c, vpbs_mlp *net, double alpha_conv, double rcut_conv, double feature_radius,
                       double sample_fraction, double tol, uint64_t seed) {
    if (!c) return VPBS_ERR_INVALID;
    c->lc_ready = false;
    if (!(alpha_conv > 0) || !(rcut_conv > 0) || !(feature_radius > 0))
        return fail(c, VPBS_ERR_INVALID, "alpha_conv, rcut_conv and feature_radius must be positive");
    if (!(alpha_conv < c->p.alpha) || !(rcut_conv > c->p.rcut))
        return fail(c, VPBS_ERR_INVALID, "the converged splitting needs alpha_conv < alpha and rcut_conv > rcut");
    if (!(sample_fraction >= 0 && sample_fraction <= 1)) return fail(c, VPBS_ERR_INVALID, "sample_fraction in [0, 1]");
    if (!(tol > 0)) return fail(c, VPBS_ERR_INVALID, "tol must be positive");
    for (int d = 0; d < 3; d++) {
        const int mw = (int)std::ceil(rcut_conv * c->Nc[d] / c->p.Ls[d] - 1e-12);
        const int mf = (int)std::ceil(feature_radius * c->Nc[d] / c->p.Ls[d] - 1e-12);
        if (2 * mw + 1 > c->Nc[d] || 2 * mw + 1 > 33)
            return fail(c, VPBS_ERR_INVALID,
                        "rcut_conv = %g spans %d cells of the cell list in direction %d (at most min(Nc, 33) = %d)",
                        rcut_conv, 2 * mw + 1, d, std::min(c->Nc[d], 33));
        if (2 * mf + 1 > c->Nc[d] || 2 * mf + 1 > 9)
            return fail(c, VPBS_ERR_INVALID,
                        "feature_radius = %g spans %d cells of the cell list in direction %d (at most min(Nc, 9))",
                        feature_radius, 2 * mf + 1, d);
    }
    if (net && (mlp_n_in(net) != LC_NIN || mlp_n_out(net) != 3))
        return fail(c, VPBS_ERR_INVALID, "the network must take %d inputs and give 3 outputs", LC_NIN);
    c->lc_net = net;
    c->lc_alpha_c = c->p.alpha;
    c->lc_alpha_f = alpha_conv;
    c->lc_rcut_f = rcut_conv;
    c->lc_rfeat = feature_radius;
    c->lc_frac = sample_fraction;
    c->lc_tol = tol;
    c->lc_seed = seed;
    c->lc_calls = 0;
    c->lc_fallbacks = 0;
    c->lc_max_err = 0.0;
    c->lc_ready = true;
    return VPBS_OK;
}

int vpbs_learned_stats(vpbs_ctx *c, int64_t *calls, int64_t *fallbacks, double *max_err) {
    if (!c) return VPBS_ERR_INVALID;
    if (calls) *calls = (int64_t)c->lc_calls;
    if (fallbacks) *fallbacks = c->lc_fallbacks;
    if (max_err) *max_err = c->lc_max_err;
    return VPBS_OK;
}

int vpbs_learned_features(vpbs_ctx *c, float *X, float *frame) {
    if (!c) return VPBS_ERR_INVALID;
    VPBS_HIP(c, hipSetDevice(c->device));
    VPBS_TRY(state_check(c, false));
    hipStream_t s = c->s_sr;
    VPBS_TRY(run_features(c, s));
    if (X) VPBS_HIP(c, hipMemcpyAsync(X, c->lc_feat, sizeof(float) * LC_NIN * (size_t)c->Np, hipMemcpyDefault, s));
    if (frame) VPBS_HIP(c, hipMemcpyAsync(frame, c->lc_frame, sizeof(float) * 9 * (size_t)c->Np, hipMemcpyDefault, s));
    VPBS_HIP(c, hipStreamSynchronize(s));
    return VPBS_OK;
}

int vpbs_learned_difference(vpbs_ctx *c, double *dv, float *T) {
    if (!c) return VPBS_ERR_INVALID;
    VPBS_HIP(c, hipSetDevice(c->device));
    VPBS_TRY(state_check(c, true));
    VPBS_TRY(lc_buffers(c));
    const int64_t Np = c->Np;
    VPBS_TRY(run_check_sr(c, Np, nullptr, c->lc_dsr, c->s_sr));
    VPBS_TRY(run_check_lr(c, Np, c->nodes, c->lc_dlr));
    VPBS_HIP(c, hipEventRecord(c->lc_ev, c->s_lr));
    VPBS_HIP(c, hipStreamWaitEvent(c->s_sr, c->lc_ev, 0));
    hipStream_t s = c->s_sr;
    if (T) {
        target_kernel<<<nblk(Np, 256), 256, 0, s>>>(Np, c->sp, c->spp, c->lc_alpha_c,
                                                    4 * M_PI / (c->p.Gamma * c->lc_alpha_c), c->lc_dsr, c->lc_dlr,
                                                    c->lc_y);
        VPBS_HIP(c, hipMemcpyAsync(T, c->lc_y, sizeof(float) * 3 * (size_t)Np, hipMemcpyDefault, s));
    }
    if (dv) {
        // dv = dsr + dlr, written into the caller's array through the dvp buffer
        VPBS_HIP(c, hipMemsetAsync(c->lc_dvp, 0, sizeof(double) * 3 * (size_t)Np, s));
        correction_exact<<<nblk(3 * Np, 256), 256, 0, s>>>(Np, nullptr, c->lc_dsr, c->lc_dlr, c->lc_dvp);
        VPBS_HIP(c, hipMemcpyAsync(dv, c->lc_dvp, sizeof(double) * 3 * (size_t)Np, hipMemcpyDefault, s));
    }
    VPBS_HIP(c, hipGetLastError());
    VPBS_HIP(c, hipStreamSynchronize(s));
    return VPBS_OK;
}

int vpbs_compute_learned(vpbs_ctx *c, double *v_out, vpbs_learned_report *rep) {
    if (!c) return VPBS_ERR_INVALID;
    if (!c->lc_ready || !c->lc_net)
        return fail(c, VPBS_ERR_STATE, "learned correction not set up with a network: call vpbs_learned_setup");
    if (!c->have_filaments)
        return fail(c, VPBS_ERR_STATE, "the learned correction needs filaments on the device (vpbs_set_filaments)");
    VPBS_HIP(c, hipSetDevice(c->device));
    VPBS_TRY(lc_buffers(c));
    auto t0 = std::chrono::steady_clock::now();
    const int64_t Np = c->Np;
    int64_t ns = c->lc_frac > 0 ? (int64_t)std::ceil(c->lc_frac * (double)Np) : 0;
    ns = std::min(ns, Np);
    c->lc_ns = ns;
    if (ns > 0) {
        // the sample, drawn first on the long-range stream; its long-range check runs there after
        // the long range, its short-range check after the cell lists (inside vpbs_compute)
        VPBS_TRY(sample_buffers(c, ns));
        draw_sample<<<nblk(ns, 256), 256, 0, c->s_lr>>>(Np, ns, c->lc_seed + c->lc_calls, c->nodes, c->lc_sidx,
                                                         c->lc_spos);
        VPBS_HIP(c, hipGetLastError());
    }
    c->lc_active = true;
    const int rc = vpbs_compute(c, VPBS_VELOCITY | VPBS_KEEP_UHAT, NAN, nullptr, nullptr);
    c->lc_active = false;
    if (rc != VPBS_OK) return rc;
    StreamDrain drain{c};
    vpbs_learned_report R{};
    R.fallback = 0;
    R.nsample = ns;
    double sums[3] = {0, 0, 0};
    if (ns > 0) {
        hipStream_t s = c->s_sr;
        const int nvp = (int)std::min<int64_t>(CHK_NV, (int64_t)nblk(Np, 256));
        vnorm_partial<<<nvp, 256, 0, s>>>(Np, c->v_out, c->lc_stats + 4);
        check_reduce<<<1, 1024, 0, s>>>(Np, ns, c->lc_sidx, c->lc_dsr, c->lc_dlr, c->lc_dvp, c->lc_stats + 4, nvp,
                                        c->lc_stats);
        VPBS_HIP(c, hipGetLastError());
        VPBS_HIP(c, hipMemcpyAsync(sums, c->lc_stats, sizeof sums, hipMemcpyDeviceToHost, s));
        VPBS_HIP(c, hipStreamSynchronize(s));
        R.err_corrected = sums[2] > 0 ? std::sqrt(sums[0] / sums[2]) : 0.0;
        R.err_uncorrected = sums[2] > 0 ? std::sqrt(sums[1] / sums[2]) : 0.0;
        if (!(R.err_corrected <= c->lc_tol)) {
            // rejected: the converged setting on every node
            R.fallback = 1;
            VPBS_TRY(run_check_sr(c, Np, nullptr, c->lc_dsr, s));
            VPBS_HIP(c, hipEventRecord(c->lc_ev, s));
            VPBS_HIP(c, hipStreamWaitEvent(c->s_lr, c->lc_ev, 0));
            VPBS_TRY(run_check_lr(c, Np, c->nodes, c->lc_dlr));
            VPBS_HIP(c, hipEventRecord(c->lc_ev, c->s_lr));
            VPBS_HIP(c, hipStreamWaitEvent(s, c->lc_ev, 0));
            correction_exact<<<nblk(3 * Np, 256), 256, 0, s>>>(Np, c->lc_dvp, c->lc_dsr, c->lc_dlr, c->v_out);
            VPBS_HIP(c, hipGetLastError());
        }
    }
    c->lc_calls++;
    c->lc_fallbacks += R.fallback;
    if (R.nsample > 0) c->lc_max_err = std::max(c->lc_max_err, R.err_corrected);
    VPBS_HIP(c, hipStreamSynchronize(c->s_sr));
    VPBS_HIP(c, hipStreamSynchronize(c->s_lr));
    auto t1 = std::chrono::steady_clock::now();
    R.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (rep) *rep = R;
    if (v_out) VPBS_HIP(c, hipMemcpy(v_out, c->v_out, sizeof(double) * 3 * (size_t)Np, hipMemcpyDeviceToHost));
    return VPBS_OK;
}

}  // extern "C"
