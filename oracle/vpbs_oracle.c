/*
 * vpbs_oracle.c — FP64 CPU restatement of VortexPasta's Ewald-split
 * Biot–Savart evaluation on filament nodes.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the parity checker for the HIP
 * library; it is linked only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py. The product path never calls it.
 *
 * Reference functions restated here (paths relative to the reference root):
 *   - short-range pair sum: src/BiotSavart/shortrange/shortrange.jl:216-273
 *     (nosimd variant), integrands shortrange/integrands.jl:2-18,
 *     weights splitting/gaussian.jl:273-279, min-image fold
 *     src/Filaments/min_distance.jl:61-67, local-segment test
 *     shortrange.jl:150-153, cut-off test r² ≤ rcut² (params.jl:107).
 *   - cell lists: src/CellLists/CellLists.jl:131-137 (cell count),
 *     :194-204 (cell index 1 + trunc((x/L)·N)), :217-220 (±M neighbours).
 *   - self-interaction removal: shortrange/local_integrals.jl:14-95.
 *   - local (LIA) term: shortrange/local_integrals.jl:106-165.
 *   - long range, exact Fourier sums: longrange/backends/exact_sum.jl:284-333.
 *   - long range, NUFFT spreading / interpolation with an ES kernel: the
 *     algorithm of NonuniformFFTs.jl (not vendored; call sites
 *     longrange/backends/nonuniformffts.jl:164,188-189,209-210), restated from
 *     its published description (Barnett et al., SIAM J. Sci. Comput. 2019).
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -shared -fPIC (see Makefile).
 * -ffp-contract=off keeps r² = (x² + y²) + z² rounded like the reference, so
 * that cut-off decisions — and hence neighbour sets — are reproducible bit
 * for bit.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FLAG_VEL 1u
#define FLAG_PSI 2u
#define FLAG_AVOID_ERF 4u  /* avoid_explicit_erf = true (include local segments) */

static const double TWO_OVER_SQRT_PI = 1.1283791670955126; /* 2/√π */

/* splitmix64 finaliser, used for an order-independent neighbour-set hash */
static inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint64_t or_pair_hash(int64_t j) { return mix64((uint64_t)j); }

/* deperiodise_separation_folded (min_distance.jl:61-67) */
static inline double fold_sep(double r, double L, double Lh) {
    r = (r >= Lh) ? r - L : r;
    r = (r < -Lh) ? r + L : r;
    return r;
}

/* number of cells per direction: floor((L·M)/rcut) (CellLists.jl:132-137) */
void or_cell_dims(double rcut, const double *Ls, int M, int32_t *Nc) {
    for (int d = 0; d < 3; d++) Nc[d] = (int32_t)floor((Ls[d] * (double)M) / rcut);
}

/* 0-based cell index of a folded coordinate: trunc((x/L)·N) (CellLists.jl:194-204) */
static inline int cell_of(double x, double L, int N) { return (int)((x / L) * (double)N); }

typedef struct {
    int32_t Nc[3];
    int64_t ncells;
    int64_t *start; /* ncells + 1 */
    int64_t *perm;  /* sources sorted by cell, stable in source index */
} celllist;

static void celllist_build(celllist *cl, const double *pts, int64_t Nq, const double *Ls) {
    const double *px = pts, *py = pts + Nq, *pz = pts + 2 * Nq;
    cl->ncells = (int64_t)cl->Nc[0] * cl->Nc[1] * cl->Nc[2];
    cl->start = calloc((size_t)cl->ncells + 1, sizeof(int64_t));
    cl->perm = malloc((size_t)(Nq > 0 ? Nq : 1) * sizeof(int64_t));
    int64_t *key = malloc((size_t)(Nq > 0 ? Nq : 1) * sizeof(int64_t));
    for (int64_t j = 0; j < Nq; j++) {
        int ix = cell_of(px[j], Ls[0], cl->Nc[0]);
        int iy = cell_of(py[j], Ls[1], cl->Nc[1]);
        int iz = cell_of(pz[j], Ls[2], cl->Nc[2]);
        key[j] = ((int64_t)iz * cl->Nc[1] + iy) * cl->Nc[0] + ix;
        cl->start[key[j] + 1]++;
    }
    for (int64_t c = 0; c < cl->ncells; c++) cl->start[c + 1] += cl->start[c];
    int64_t *fill = malloc((size_t)cl->ncells * sizeof(int64_t));
    memcpy(fill, cl->start, (size_t)cl->ncells * sizeof(int64_t));
    for (int64_t j = 0; j < Nq; j++) cl->perm[fill[key[j]]++] = j;
    free(fill);
    free(key);
}

static void celllist_free(celllist *cl) {
    free(cl->start);
    free(cl->perm);
}

/* Accumulate one source j into node i's sums (shortrange.jl:240-268). */
static inline int pair_term(const double x[3], const double *pts, const double *chg, int64_t Nq, int64_t j,
                            const double *Ls, const double *Lhs, double rcut2, double alpha, unsigned flags,
                            double v[3], double psi[3]) {
    double r[3];
    for (int d = 0; d < 3; d++) r[d] = fold_sep(x[d] - pts[d * Nq + j], Ls[d], Lhs[d]);
    double r2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
    if (!(r2 <= rcut2)) return 0;
    double rr = sqrt(r2);
    double rinv = 1.0 / rr;
    double ar = alpha * rr;
    double a = erfc(ar);
    double b = TWO_OVER_SQRT_PI * ar * exp(-(ar * ar));
    double q[3] = {chg[j], chg[Nq + j], chg[2 * Nq + j]};
    if (flags & FLAG_VEL) {
        double f = (a + b) * (rinv * rinv * rinv);
        v[0] += f * (q[1] * r[2] - q[2] * r[1]);
        v[1] += f * (q[2] * r[0] - q[0] * r[2]);
        v[2] += f * (q[0] * r[1] - q[1] * r[0]);
    }
    if (flags & FLAG_PSI) {
        double f = a * rinv;
        psi[0] += f * q[0];
        psi[1] += f * q[1];
        psi[2] += f * q[2];
    }
    return 1;
}

/*
 * Short-range pair sum (add_pair_interactions!). Results are ADDED to v/psi
 * (component-major, 3×Np), multiplied by Γ/4π. With use_celllist = 0 every
 * source is visited (the NaiveShortRangeBackend, backends/naive.jl:217-229).
 * counts/hashes (optional) receive, per node, the number of accepted pairs and
 * the wrapping sum of mix64(j) over accepted j. Only nodes in [i_begin, i_end)
 * are processed (i_end < 0: all), which lets a CPU timing run take a sample.
 */
void or_shortrange(int64_t Np, const double *nodes, const int32_t *idx_prev, int64_t Nq, const double *pts,
                   const double *chg, int nquad, const double *Ls, double alpha, double rcut, double Gamma,
                   int nsubdiv, int use_celllist, unsigned flags, double *v, double *psi, int64_t *counts,
                   uint64_t *hashes, int64_t i_begin, int64_t i_end) {
    const double Lhs[3] = {Ls[0] / 2, Ls[1] / 2, Ls[2] / 2};
    const double rcut2 = rcut * rcut;
    const double pref = Gamma / (4 * M_PI);
    const int include_local = (flags & FLAG_AVOID_ERF) != 0;
    celllist cl;
    memset(&cl, 0, sizeof cl);
    const int M = nsubdiv;
    if (use_celllist) {
        or_cell_dims(rcut, Ls, M, cl.Nc);
        celllist_build(&cl, pts, Nq, Ls);
    }
    if (i_end < 0 || i_end > Np) i_end = Np;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = i_begin; i < i_end; i++) {
        double x[3] = {nodes[i], nodes[Np + i], nodes[2 * Np + i]};
        double vi[3] = {0, 0, 0}, pi[3] = {0, 0, 0};
        int64_t cnt = 0;
        uint64_t h = 0;
        const int64_t jr0 = (int64_t)nquad * i, jl0 = (int64_t)nquad * idx_prev[i];
#define VISIT(j)                                                                                           \
    do {                                                                                                   \
        int64_t jj = (j);                                                                                  \
        if (!include_local && ((jj >= jr0 && jj < jr0 + nquad) || (jj >= jl0 && jj < jl0 + nquad))) break; \
        if (pair_term(x, pts, chg, Nq, jj, Ls, Lhs, rcut2, alpha, flags, vi, pi)) {                       \
            cnt++;                                                                                         \
            h += mix64((uint64_t)jj);                                                                      \
        }                                                                                                  \
    } while (0)
        if (use_celllist) {
            int c0[3];
            for (int d = 0; d < 3; d++) c0[d] = cell_of(x[d], Ls[d], cl.Nc[d]);
            for (int dz = -M; dz <= M; dz++) {
                int cz = ((c0[2] + dz) % cl.Nc[2] + cl.Nc[2]) % cl.Nc[2];
                for (int dy = -M; dy <= M; dy++) {
                    int cy = ((c0[1] + dy) % cl.Nc[1] + cl.Nc[1]) % cl.Nc[1];
                    for (int dx = -M; dx <= M; dx++) {
                        int cx = ((c0[0] + dx) % cl.Nc[0] + cl.Nc[0]) % cl.Nc[0];
                        int64_t c = ((int64_t)cz * cl.Nc[1] + cy) * cl.Nc[0] + cx;
                        for (int64_t s = cl.start[c]; s < cl.start[c + 1]; s++) VISIT(cl.perm[s]);
                    }
                }
            }
        } else {
            for (int64_t j = 0; j < Nq; j++) VISIT(j);
        }
#undef VISIT
        if (flags & FLAG_VEL)
            for (int d = 0; d < 3; d++) v[d * Np + i] += pref * vi[d];
        if (flags & FLAG_PSI)
            for (int d = 0; d < 3; d++) psi[d * Np + i] += pref * pi[d];
        if (counts) counts[i] = cnt;
        if (hashes) hashes[i] = h;
    }
    if (use_celllist) celllist_free(&cl);
}

/*
 * Open-domain (NoSplitting) direct sum over all sources except the two local
 * segments: weights (1, 0) (splitting.jl:332-333), no folding. Used to pin
 * the integrands and local terms against vortex-ring closed forms.
 */
void or_direct_open(int64_t Np, const double *nodes, const int32_t *idx_prev, int64_t Nq, const double *pts,
                    const double *chg, int nquad, double Gamma, unsigned flags, double *v, double *psi) {
    const double pref = Gamma / (4 * M_PI);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Np; i++) {
        double x[3] = {nodes[i], nodes[Np + i], nodes[2 * Np + i]};
        double vi[3] = {0, 0, 0}, pi[3] = {0, 0, 0};
        const int64_t jr0 = (int64_t)nquad * i, jl0 = (int64_t)nquad * idx_prev[i];
        for (int64_t j = 0; j < Nq; j++) {
            if ((j >= jr0 && j < jr0 + nquad) || (j >= jl0 && j < jl0 + nquad)) continue;
            double r[3];
            for (int d = 0; d < 3; d++) r[d] = x[d] - pts[d * Nq + j];
            double r2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
            double rinv = 1.0 / sqrt(r2);
            double q[3] = {chg[j], chg[Nq + j], chg[2 * Nq + j]};
            if (flags & FLAG_VEL) {
                double f = rinv * rinv * rinv;
                vi[0] += f * (q[1] * r[2] - q[2] * r[1]);
                vi[1] += f * (q[2] * r[0] - q[0] * r[2]);
                vi[2] += f * (q[0] * r[1] - q[1] * r[0]);
            }
            if (flags & FLAG_PSI)
                for (int d = 0; d < 3; d++) pi[d] += rinv * q[d];
        }
        if (flags & FLAG_VEL)
            for (int d = 0; d < 3; d++) v[d * Np + i] += pref * vi[d];
        if (flags & FLAG_PSI)
            for (int d = 0; d < 3; d++) psi[d * Np + i] += pref * pi[d];
    }
}

/*
 * remove_self_interaction! (local_integrals.jl:14-95): subtract, over the
 * quadrature points of the two segments adjacent to each node, the full
 * integrand (avoid_explicit_erf) or its long-range (erf) part.
 */
void or_self_interaction(int64_t Np, const double *nodes, const int32_t *idx_prev, int64_t Nq, const double *pts,
                         const double *chg, int nquad, const double *Ls, double alpha, double Gamma,
                         unsigned flags, double *v, double *psi) {
    const double Lhs[3] = {Ls[0] / 2, Ls[1] / 2, Ls[2] / 2};
    const double pref = Gamma / (4 * M_PI);
    const int full = (flags & FLAG_AVOID_ERF) != 0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Np; i++) {
        double x[3] = {nodes[i], nodes[Np + i], nodes[2 * Np + i]};
        for (int side = 0; side < 2; side++) {
            int64_t j0 = (int64_t)nquad * (side == 0 ? idx_prev[i] : i);
            for (int k = 0; k < nquad; k++) {
                int64_t j = j0 + k;
                double r[3];
                for (int d = 0; d < 3; d++) r[d] = fold_sep(x[d] - pts[d * Nq + j], Ls[d], Lhs[d]);
                double r2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
                double rr = sqrt(r2);
                double rinv = 1.0 / rr;
                double a = 1.0, b = 0.0;
                if (!full) {
                    double ar = alpha * rr;
                    a = erf(ar);
                    b = TWO_OVER_SQRT_PI * ar * exp(-(ar * ar));
                }
                double q[3] = {chg[j], chg[Nq + j], chg[2 * Nq + j]};
                if (flags & FLAG_VEL) {
                    double f = (a - b) * (rinv * rinv * rinv);
                    v[i] -= pref * (f * (q[1] * r[2] - q[2] * r[1]));
                    v[Np + i] -= pref * (f * (q[2] * r[0] - q[0] * r[2]));
                    v[2 * Np + i] -= pref * (f * (q[0] * r[1] - q[1] * r[0]));
                }
                if (flags & FLAG_PSI) {
                    double f = a * rinv;
                    for (int d = 0; d < 3; d++) psi[d * Np + i] -= pref * (f * q[d]);
                }
            }
        }
    }
}

/*
 * compute_local_term! (local_integrals.jl:106-165). REPLACES v/psi.
 * cutoff_v/cutoff_psi are exp(Δ)a/2 and exp(Δ-1)a/2 unless an
 * integration_cutoff is given (both then equal it).
 */
void or_local_term(int64_t Np, const double *sp, const double *spp, const double *dlm, const double *dlp,
                   double Gamma, double cutoff_v, double cutoff_psi, unsigned flags, double *v, double *psi) {
    const double pref = Gamma / (4 * M_PI);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Np; i++) {
        double delta = sqrt(dlm[i] * dlp[i]);
        double s1[3] = {sp[i], sp[Np + i], sp[2 * Np + i]};
        double s2[3] = {spp[i], spp[Np + i], spp[2 * Np + i]};
        double mu2 = (s1[0] * s1[0] + s1[1] * s1[1]) + s1[2] * s1[2];
        double muinv = 1.0 / sqrt(mu2);
        if (flags & FLAG_VEL) {
            double m3 = muinv * muinv * muinv;
            double bx = (s1[1] * s2[2] - s1[2] * s2[1]) * m3;
            double by = (s1[2] * s2[0] - s1[0] * s2[2]) * m3;
            double bz = (s1[0] * s2[1] - s1[1] * s2[0]) * m3;
            double lg = log(delta / cutoff_v);
            v[i] = pref * (lg * bx);
            v[Np + i] = pref * (lg * by);
            v[2 * Np + i] = pref * (lg * bz);
        }
        if (flags & FLAG_PSI) {
            double lg = 2 * log(delta / cutoff_psi);
            for (int d = 0; d < 3; d++) psi[d * Np + i] = pref * (lg * (s1[d] * muinv));
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Long range                                                                */

/* Retained Fourier modes per direction, after the reference's Hermitian
 * clean-up (longrange.jl:556-583): kx ∈ [0, N÷2 - 1]; ky, kz ∈ [-K, K] with
 * K = N/2 - 1 (N even) or (N - 1)/2 (N odd). */
static void mode_limits(const int32_t *Ns, int *kxmax, int *kymax, int *kzmax) {
    *kxmax = Ns[0] / 2 - 1;
    *kymax = (Ns[1] % 2 == 0) ? Ns[1] / 2 - 1 : (Ns[1] - 1) / 2;
    *kzmax = (Ns[2] % 2 == 0) ? Ns[2] / 2 - 1 : (Ns[2] - 1) / 2;
}

/*
 * ExactSum type 1 (exact_sum.jl:284-301): uhat(k) = pref · Σ_j q_j e^{-i k·s_j}
 * on the r2c grid of shape (Nz, Ny, Nx/2+1) (x fastest), complex interleaved,
 * 3 components one after the other. Non-retained modes are set to zero.
 */
void or_exact_type1(int64_t Nq, const double *pts, const double *chg, const double *Ls, const int32_t *Ns,
                    double pref, double *uhat) {
    const int Nkx = Ns[0] / 2 + 1, Ny = Ns[1], Nz = Ns[2];
    const int64_t nmodes = (int64_t)Nkx * Ny * Nz;
    int kxmax, kymax, kzmax;
    mode_limits(Ns, &kxmax, &kymax, &kzmax);
    memset(uhat, 0, sizeof(double) * 2 * 3 * (size_t)nmodes);
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < Nz; iz++) {
        for (int iy = 0; iy < Ny; iy++) {
            int kz = iz < (Nz + 1) / 2 ? iz : iz - Nz; /* fftfreq ordering */
            int ky = iy < (Ny + 1) / 2 ? iy : iy - Ny;
            if (kz < -kzmax || kz > kzmax || ky < -kymax || ky > kymax) continue;
            double kyp = 2 * M_PI * ky / Ls[1], kzp = 2 * M_PI * kz / Ls[2];
            for (int ix = 0; ix <= kxmax; ix++) {
                double kxp = 2 * M_PI * ix / Ls[0];
                double acc[6] = {0, 0, 0, 0, 0, 0};
                for (int64_t j = 0; j < Nq; j++) {
                    double ph = -(kxp * pts[j] + kyp * pts[Nq + j] + kzp * pts[2 * Nq + j]);
                    double c = cos(ph), s = sin(ph);
                    for (int d = 0; d < 3; d++) {
                        acc[2 * d] += chg[d * Nq + j] * c;
                        acc[2 * d + 1] += chg[d * Nq + j] * s;
                    }
                }
                int64_t m = ((int64_t)iz * Ny + iy) * Nkx + ix;
                for (int d = 0; d < 3; d++) {
                    uhat[2 * (d * nmodes + m)] = pref * acc[2 * d];
                    uhat[2 * (d * nmodes + m) + 1] = pref * acc[2 * d + 1];
                }
            }
        }
    }
}

/*
 * ExactSum type 2 (exact_sum.jl:303-333): out(x_i) = Σ_k Re(û(k) e^{ik·x_i})
 * counting kx > 0 modes twice (Hermitian symmetry). ``uhat`` as in type 1.
 */
void or_exact_type2(int64_t Np, const double *nodes, const double *Ls, const int32_t *Ns, const double *uhat,
                    double *out) {
    const int Nkx = Ns[0] / 2 + 1, Ny = Ns[1], Nz = Ns[2];
    const int64_t nmodes = (int64_t)Nkx * Ny * Nz;
    int kxmax, kymax, kzmax;
    mode_limits(Ns, &kxmax, &kymax, &kzmax);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Np; i++) {
        double acc[3] = {0, 0, 0};
        for (int iz = 0; iz < Nz; iz++) {
            int kz = iz < (Nz + 1) / 2 ? iz : iz - Nz;
            if (kz < -kzmax || kz > kzmax) continue;
            for (int iy = 0; iy < Ny; iy++) {
                int ky = iy < (Ny + 1) / 2 ? iy : iy - Ny;
                if (ky < -kymax || ky > kymax) continue;
                for (int ix = 0; ix <= kxmax; ix++) {
                    double ph = 2 * M_PI * ix / Ls[0] * nodes[i] + 2 * M_PI * ky / Ls[1] * nodes[Np + i] +
                                2 * M_PI * kz / Ls[2] * nodes[2 * Np + i];
                    double c = cos(ph), s = sin(ph);
                    double w = (ix == 0) ? 1.0 : 2.0;
                    int64_t m = ((int64_t)iz * Ny + iy) * Nkx + ix;
                    for (int d = 0; d < 3; d++) {
                        double re = uhat[2 * (d * nmodes + m)], im = uhat[2 * (d * nmodes + m) + 1];
                        acc[d] += w * (re * c - im * s);
                    }
                }
            }
        }
        for (int d = 0; d < 3; d++) out[d * Np + i] = acc[d];
    }
}

/* ES kernel φ(z) = exp(β(√(1 - z²) - 1)) for |z| ≤ 1 */
static inline double es_kernel(double z, double beta) {
    double s = 1.0 - z * z;
    return s < 0 ? 0.0 : exp(beta * (sqrt(s) - 1.0));
}

static inline void es_weights(double u, int M, double beta, int64_t *i0, double *w) {
    double fl = floor(u);
    *i0 = (int64_t)fl - M + 1;
    for (int k = 0; k < 2 * M; k++) w[k] = es_kernel(((double)(*i0 + k) - u) / (double)M, beta);
}

/*
 * NUFFT spreading on the oversampled grid G (shape Gz×Gy×Gx, x fastest, 3
 * components one after the other): grid[n] += q_j · Π_d φ(n_d - u_{j,d})
 * with u = x·G/L, periodically wrapped. ``grid`` must be zeroed by the caller.
 * Threads own disjoint sets of z-planes and visit the points in index order,
 * so every grid value is summed in the same (deterministic) order.
 */
void or_spread(int64_t Nq, const double *pts, const double *chg, const double *Ls, const int32_t *G, int M,
               double beta, double *grid) {
    const int64_t gsz = (int64_t)G[0] * G[1] * G[2];
    const int W = 2 * M;
    double *wts = malloc(sizeof(double) * (size_t)(Nq > 0 ? Nq : 1) * 3 * W);
    int32_t *idx = malloc(sizeof(int32_t) * (size_t)(Nq > 0 ? Nq : 1) * 3 * W);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < Nq; j++) {
        for (int d = 0; d < 3; d++) {
            int64_t i0;
            es_weights(pts[d * Nq + j] * ((double)G[d] / Ls[d]), M, beta, &i0, wts + (j * 3 + d) * W);
            for (int k = 0; k < W; k++) idx[(j * 3 + d) * W + k] = (int32_t)(((i0 + k) % G[d] + G[d]) % G[d]);
        }
    }
#pragma omp parallel
    {
#ifdef _OPENMP
        int nt = omp_get_num_threads(), tid = omp_get_thread_num();
#else
        int nt = 1, tid = 0;
#endif
        for (int64_t j = 0; j < Nq; j++) {
            const int32_t *iz = idx + (j * 3 + 2) * W;
            const int32_t *iy = idx + (j * 3 + 1) * W;
            const int32_t *ix = idx + (j * 3 + 0) * W;
            const double *wz = wts + (j * 3 + 2) * W;
            const double *wy = wts + (j * 3 + 1) * W;
            const double *wx = wts + (j * 3 + 0) * W;
            for (int c = 0; c < W; c++) {
                if (iz[c] % nt != tid) continue; /* plane owned by another thread */
                for (int comp = 0; comp < 3; comp++) {
                    double qz = chg[comp * Nq + j] * wz[c];
                    double *g = grid + comp * gsz + (int64_t)iz[c] * G[1] * G[0];
                    for (int b = 0; b < W; b++) {
                        double qzy = qz * wy[b];
                        double *row = g + (int64_t)iy[b] * G[0];
                        for (int a = 0; a < W; a++) row[ix[a]] += qzy * wx[a];
                    }
                }
            }
        }
    }
    free(wts);
    free(idx);
}

/* NUFFT interpolation: out_i = Σ_n Π_d φ(u_{i,d} - n_d) grid[n] (3 components). */
void or_interp(int64_t Np, const double *nodes, const double *Ls, const int32_t *G, int M, double beta,
               const double *grid, double *out) {
    const int64_t gsz = (int64_t)G[0] * G[1] * G[2];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Np; i++) {
        double wx[32], wy[32], wz[32];
        int64_t ix0, iy0, iz0;
        es_weights(nodes[i] * ((double)G[0] / Ls[0]), M, beta, &ix0, wx);
        es_weights(nodes[Np + i] * ((double)G[1] / Ls[1]), M, beta, &iy0, wy);
        es_weights(nodes[2 * Np + i] * ((double)G[2] / Ls[2]), M, beta, &iz0, wz);
        for (int comp = 0; comp < 3; comp++) {
            const double *g = grid + comp * gsz;
            double acc = 0;
            for (int c = 0; c < 2 * M; c++) {
                int64_t iz = ((iz0 + c) % G[2] + G[2]) % G[2];
                for (int b = 0; b < 2 * M; b++) {
                    int64_t iy = ((iy0 + b) % G[1] + G[1]) % G[1];
                    const double *row = g + (iz * G[1] + iy) * G[0];
                    double accx = 0;
                    for (int a = 0; a < 2 * M; a++) {
                        int64_t ix = ((ix0 + a) % G[0] + G[0]) % G[0];
                        accx += wx[a] * row[ix];
                    }
                    acc += wz[c] * wy[b] * accx;
                }
            }
            out[comp * Np + i] = acc;
        }
    }
}

/*
 * add_local_integrals! (local_integrals.jl:250-296): for each node i, add the
 * full Biot–Savart integrand over nsub sub-segment quadrature points that the
 * caller built (points spts[3][nsub*Np], charges schg, node-major: node i owns
 * entries nsub*i .. nsub*i+nsub-1). periodic != 0: min-image separations.
 */
void or_local_integrals(int64_t Np, const double *nodes, int nsub, const double *spts, const double *schg,
                        const double *Ls, int periodic, double Gamma, unsigned flags, double *v, double *psi) {
    const double pref = Gamma / (4 * M_PI);
    const int64_t n = (int64_t)nsub * Np;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < Np; i++) {
        double x[3] = {nodes[i], nodes[Np + i], nodes[2 * Np + i]};
        double vi[3] = {0, 0, 0}, pi[3] = {0, 0, 0};
        for (int k = 0; k < nsub; k++) {
            int64_t j = (int64_t)nsub * i + k;
            double r[3];
            for (int d = 0; d < 3; d++) {
                r[d] = x[d] - spts[d * n + j];
                if (periodic) r[d] = fold_sep(r[d], Ls[d], Ls[d] / 2);
            }
            double r2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
            double rinv = 1.0 / sqrt(r2);
            double q[3] = {schg[j], schg[n + j], schg[2 * n + j]};
            double f = rinv * rinv * rinv;
            vi[0] += f * (q[1] * r[2] - q[2] * r[1]);
            vi[1] += f * (q[2] * r[0] - q[0] * r[2]);
            vi[2] += f * (q[0] * r[1] - q[1] * r[0]);
            for (int d = 0; d < 3; d++) pi[d] += rinv * q[d];
        }
        if (flags & FLAG_VEL)
            for (int d = 0; d < 3; d++) v[d * Np + i] += pref * vi[d];
        if (flags & FLAG_PSI)
            for (int d = 0; d < 3; d++) psi[d * Np + i] += pref * pi[d];
    }
}
