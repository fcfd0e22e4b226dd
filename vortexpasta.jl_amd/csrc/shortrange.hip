// Short-range part of the Ewald-split Biot–Savart sum on gfx950.
//
//   local_term_kernel   LIA term, FP64 (shortrange/local_integrals.jl:106-165)
//   cell keys + sort    device cell list (CellLists.jl:131-137,194-204), built by a
//                       deterministic counting sort instead of an atomic linked list
//   pair_kernel         erfc/Gaussian pair sum over the (2M+1)³ neighbour cells
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

#include <cmath>

namespace vpbs {

namespace {

constexpr double TWO_OVER_SQRT_PI = 1.1283791670955126;

__device__ inline double fold_sep(double r, double L, double Lh) {
    r = (r >= Lh) ? r - L : r;
    r = (r < -Lh) ? r + L : r;
    return r;
}

// r² rounded exactly as (x² + y²) + z² (no contraction), as in the reference.
__device__ inline double dist2_nocontract(double a, double b, double c) {
#pragma clang fp contract(off)
    double s = a * a + b * b;
    return s + c * c;
}

__device__ inline int cell_index(double x, double L, int N) { return (int)((x / L) * (double)N); }

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

__global__ void cell_keys(int64_t n, const double *__restrict__ pos, Box box, int3 Nc, int32_t *__restrict__ key) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int ix = cell_index(pos[j], box.L[0], Nc.x);
    int iy = cell_index(pos[n + j], box.L[1], Nc.y);
    int iz = cell_index(pos[2 * n + j], box.L[2], Nc.z);
    // Points folded into [0, L) always give indices in [0, N); clamp defensively.
    ix = min(max(ix, 0), Nc.x - 1);
    iy = min(max(iy, 0), Nc.y - 1);
    iz = min(max(iz, 0), Nc.z - 1);
    key[j] = (iz * Nc.y + iy) * Nc.x + ix;
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
    double rcut2;
    CT alpha;
    double pref;
    int include_local;
    double *v;
    double *psi;
    int64_t *counts;
    uint64_t *hashes;
};

__device__ inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

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

__device__ inline double erfc_t(double x) { return erfc(x); }
__device__ inline float erfc_t(float x) { return erfcf(x); }
__device__ inline double exp_t(double x) { return exp(x); }
__device__ inline float exp_t(float x) { return expf(x); }
__device__ inline double sqrt_t(double x) { return sqrt(x); }
__device__ inline float sqrt_t(float x) { return sqrtf(x); }

template <typename CT, bool VEL, bool PSI, bool STATS>
__global__ __launch_bounds__(256) void pair_kernel(PairArgs<CT> a) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.Np) return;
    const int64_t Np = a.Np, Nq = a.Nq;
    const int32_t i = a.tperm[k];
    const double x = a.nodes[i], y = a.nodes[Np + i], z = a.nodes[2 * Np + i];
    const int cx = min(max(cell_index(x, a.box.L[0], a.Nc.x), 0), a.Nc.x - 1);
    const int cy = min(max(cell_index(y, a.box.L[1], a.Nc.y), 0), a.Nc.y - 1);
    const int cz = min(max(cell_index(z, a.box.L[2], a.Nc.z), 0), a.Nc.z - 1);
    const int64_t jr0 = (int64_t)a.nquad * i, jl0 = (int64_t)a.nquad * a.idx_prev[i];
    const int M = a.M;
    Acc3<CT> av, ap;
    int64_t cnt = 0;
    uint64_t h = 0;
    const CT two_over_sqrt_pi = (CT)TWO_OVER_SQRT_PI;
    for (int dz = -M; dz <= M; dz++) {
        int zz = cz + dz;
        zz += (zz < 0) ? a.Nc.z : 0;
        zz -= (zz >= a.Nc.z) ? a.Nc.z : 0;
        for (int dy = -M; dy <= M; dy++) {
            int yy = cy + dy;
            yy += (yy < 0) ? a.Nc.y : 0;
            yy -= (yy >= a.Nc.y) ? a.Nc.y : 0;
            const int64_t row = ((int64_t)zz * a.Nc.y + yy) * a.Nc.x;
            int lo = cx - M, hi = cx + M;
            // up to two contiguous runs of cells along x (periodic wrap)
            int r0a, r0b, r1a = 0, r1b = -1;
            if (lo < 0) {
                r0a = lo + a.Nc.x;
                r0b = a.Nc.x - 1;
                r1a = 0;
                r1b = hi;
            } else if (hi >= a.Nc.x) {
                r0a = lo;
                r0b = a.Nc.x - 1;
                r1a = 0;
                r1b = hi - a.Nc.x;
            } else {
                r0a = lo;
                r0b = hi;
            }
            for (int run = 0; run < 2; run++) {
                int ca = run ? r1a : r0a, cb = run ? r1b : r0b;
                if (cb < ca) continue;
                const int32_t s0 = a.cstart[row + ca], s1 = a.cstart[row + cb + 1];
                for (int32_t s = s0; s < s1; s++) {
                    const int32_t j = a.sidx[s];
                    if (!a.include_local &&
                        ((j >= jr0 && j < jr0 + a.nquad) || (j >= jl0 && j < jl0 + a.nquad)))
                        continue;
                    const double rx = fold_sep(x - a.spos[s], a.box.L[0], a.box.Lh[0]);
                    const double ry = fold_sep(y - a.spos[Nq + s], a.box.L[1], a.box.Lh[1]);
                    const double rz = fold_sep(z - a.spos[2 * Nq + s], a.box.L[2], a.box.Lh[2]);
                    const double r2 = dist2_nocontract(rx, ry, rz);
                    if (!(r2 <= a.rcut2)) continue;
                    if (STATS) {
                        cnt++;
                        h += mix64((uint64_t)j);
                    }
                    const CT r = sqrt_t((CT)r2);
                    const CT rinv = (CT)1 / r;
                    const CT ar = a.alpha * r;
                    const CT wa = erfc_t(ar);
                    const CT qx = a.schg[s], qy = a.schg[Nq + s], qz = a.schg[2 * Nq + s];
                    if (VEL) {
                        const CT wb = two_over_sqrt_pi * ar * exp_t(-(ar * ar));
                        const CT f = (wa + wb) * (rinv * rinv * rinv);
                        const CT crx = (CT)rx, cry = (CT)ry, crz = (CT)rz;
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
                }
            }
        }
    }
    if (VEL) {
        a.v[i] += a.pref * av.get(0);
        a.v[Np + i] += a.pref * av.get(1);
        a.v[2 * Np + i] += a.pref * av.get(2);
    }
    if (PSI) {
        a.psi[i] += a.pref * ap.get(0);
        a.psi[Np + i] += a.pref * ap.get(1);
        a.psi[2 * Np + i] += a.pref * ap.get(2);
    }
    if (STATS) {
        a.counts[i] = cnt;
        a.hashes[i] = h;
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
void launch_pair(const PairArgs<CT> &a, bool stats, hipStream_t s) {
    const unsigned nb = nblocks(a.Np, 256);
    if (stats)
        pair_kernel<CT, VEL, PSI, true><<<nb, 256, 0, s>>>(a);
    else
        pair_kernel<CT, VEL, PSI, false><<<nb, 256, 0, s>>>(a);
}

template <typename CT>
int run_pairs(vpbs_ctx *c, bool vel, bool psi, bool stats, bool include_local) {
    const int64_t Np = c->Np, Nq = c->Nq;
    hipStream_t s = c->s_sr;
    const int3 Nc = make_int3(c->Nc[0], c->Nc[1], c->Nc[2]);
    stage_begin(c, ST_CELL, s);
    cell_keys<<<nblocks(Nq, 256), 256, 0, s>>>(Nq, c->points, c->box, Nc, c->src_key);
    VPBS_TRY(counting_sort(c, c->sort_sr, c->src_key, Nq, c->ncells, c->src_start, c->src_perm, s));
    gather_sources<CT><<<nblocks(Nq, 256), 256, 0, s>>>(Nq, c->src_perm, c->points, c->charges, c->src_pos,
                                                          (CT *)c->src_chg, c->src_idx);
    cell_keys<<<nblocks(Np, 256), 256, 0, s>>>(Np, c->nodes, c->box, Nc, c->tgt_key);
    VPBS_TRY(counting_sort(c, c->sort_tg, c->tgt_key, Np, c->ncells, c->tgt_start, c->tgt_perm, s));
    stage_end(c, ST_CELL, s);

    PairArgs<CT> a;
    a.Np = Np;
    a.Nq = Nq;
    a.tperm = c->tgt_perm;
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
    a.rcut2 = c->p.rcut * c->p.rcut;  // params.jl:107
    a.alpha = (CT)c->p.alpha;
    a.pref = c->p.Gamma / (4 * M_PI);
    a.include_local = include_local ? 1 : 0;
    a.v = c->v_sr;
    a.psi = c->psi_sr;
    a.counts = c->stat_counts;
    a.hashes = c->stat_hashes;
    stage_begin(c, ST_PAIR, s);
    if (vel && psi)
        launch_pair<CT, true, true>(a, stats, s);
    else if (vel)
        launch_pair<CT, true, false>(a, stats, s);
    else
        launch_pair<CT, false, true>(a, stats, s);
    stage_end(c, ST_PAIR, s);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

}  // namespace

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
