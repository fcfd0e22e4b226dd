// Device-side point data: the work of add_point_charges!
// (src/BiotSavart/pointdata_impl.jl:121-188) on the GPU, from the filaments'
// periodic quintic splines (src/Filaments/spline/spline.jl:116-319).
//
// Inputs (uploaded by vpbs_set_filaments): node positions X, knots t_i
// (t = 0 at each filament's first node), knot periods T, end-to-end offsets
// Xoff and the B-spline coefficients c of the periodised curve
// Y(t) = X(t) - (t/T)·Xoff (spline.jl:163-193). The derivative coefficients
//   d_j = 5(c_j - c_{j-1})/(t_{j+5} - t_j),  e_j = 4(d_j - d_{j-1})/(t_{j+4} - t_j)
// (spline_derivative!, src/Filaments/spline/spline.jl:298-319) give S′ and S″
// as order-5 and order-4 splines. One thread per node then
// writes the folded node, idx_prev, s′, s″, δ± (Gauss–Legendre integral of |S′|
// over the LIA fraction of the two adjacent segments, pointdata_impl.jl:133-143)
// and the nquad quadrature points (folded) and charges w·Δt·S′(ζ) of its right
// segment (pointdata_impl.jl:173-183), in the layout of vpbs_set_points.
#include "vpbs_internal.hpp"

#include <cmath>
#include <vector>

namespace vpbs {

// No FMA contraction in this file: the point data comes out bit-identical to a
// plain IEEE evaluation of the same formulas (numpy's). This matters beyond
// rounding: a quadrature point lying exactly on a NUFFT grid plane (a ring in
// the plane z = L/2, say) selects its 2m-point ES stencil by floor(x·G/L), and
// a 1-ulp change there moves a weight φ(±1) = e^{-β} (≈ 1e-7 at m = 4) from one
// end of the stencil to the other.
#pragma clang fp contract(off)

namespace {

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct Fil {
    const int32_t *fid;    // [Np] filament of each node
    const int64_t *off;    // [Nf] first node
    const int32_t *N;      // [Nf] nodes per filament
    const double *T;       // [Nf] knot period
    const double *Xoff;    // [3*Nf] end-to-end offset
    const double *ts;      // [Np] knots
    int64_t Np, Nf;
};

__device__ inline double knot(const Fil &f, int64_t off, int N, double T, int j) {
    int q = j / N, r = j - q * N;
    if (r < 0) {
        r += N;
        q -= 1;
    }
    return f.ts[off + r] + q * T;
}

// Non-zero order-K B-splines at t ∈ [t_i, t_{i+1}): B[r] = B_{i-K+1+r,K}(t).
template <int K>
__device__ inline void basis(const Fil &f, int64_t off, int N, double T, int i, double t, double *B) {
    double left[K], right[K];
    B[0] = 1.0;
#pragma unroll
    for (int j = 1; j < K; j++) {
        left[j] = t - knot(f, off, N, T, i + 1 - j);
        right[j] = knot(f, off, N, T, i + j) - t;
        double saved = 0.0;
#pragma unroll
        for (int r = 0; r < j; r++) {
            const double temp = B[r] / (right[r + 1] + left[j - r]);
            B[r] = saved + right[r + 1] * temp;
            saved = left[j - r] * temp;
        }
        B[j] = saved;
    }
}

// S^(deriv)(t) on segment `seg` at local coordinate ζ, coefficient array `co`
// (c for deriv 0, d for 1, e for 2), with the periodisation undone.
template <int DERIV>
__device__ inline void eval_spline(const Fil &f, int fi, int seg, double zeta, const double *co, double *out) {
    constexpr int K = 6 - DERIV;
    const int64_t off = f.off[fi];
    const int N = f.N[fi];
    const double T = f.T[fi];
    const double ti = knot(f, off, N, T, seg), ti1 = knot(f, off, N, T, seg + 1);
    const double t = (1 - zeta) * ti + zeta * ti1;
    double B[K];
    basis<K>(f, off, N, T, seg, t, B);
    double s[3] = {0, 0, 0};
#pragma unroll
    for (int r = 0; r < K; r++) {
        int j = seg - K + 1 + r;
        j %= N;
        j += (j < 0) ? N : 0;
        const int64_t idx = off + j;
#pragma unroll
        for (int d = 0; d < 3; d++) s[d] += B[r] * co[d * f.Np + idx];
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double xo = f.Xoff[d * f.Nf + fi];
        if (DERIV == 0) s[d] += (t / T) * xo;
        if (DERIV == 1) s[d] += xo / T;
        out[d] = s[d];
    }
}

// fold_coordinates_periodic (min_distance.jl:5-20): the x < 0 loop first, so
// that a tiny negative value ends at 0, never at L. Values more than 64
// periods away are first reduced by the exact fmod (the loops alone would not
// end for |x| ≫ L); non-finite values become NaN, which the folded-range check
// after the build reports as an error.
__device__ inline double fold(double x, double L) {
    if (!isfinite(x)) return __builtin_nan("");
    if (fabs(x) > 64.0 * L) x = fmod(x, L);
    while (x < 0) x += L;
    while (x >= L) x -= L;
    return x;
}

__global__ void derivative_coefs(Fil f, const double *__restrict__ c, double *__restrict__ dco) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= f.Np) return;
    const int fi = f.fid[i];
    const int64_t off = f.off[fi];
    const int N = f.N[fi];
    const double T = f.T[fi];
    const int l = (int)(i - off);
    const int64_t ip = off + (l == 0 ? N - 1 : l - 1);
    const double den = knot(f, off, N, T, l + 5) - knot(f, off, N, T, l);
#pragma unroll
    for (int d = 0; d < 3; d++) dco[d * f.Np + i] = 5.0 * (c[d * f.Np + i] - c[d * f.Np + ip]) / den;
}

__global__ void second_derivative_coefs(Fil f, const double *__restrict__ dco, double *__restrict__ eco) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= f.Np) return;
    const int fi = f.fid[i];
    const int64_t off = f.off[fi];
    const int N = f.N[fi];
    const double T = f.T[fi];
    const int l = (int)(i - off);
    const int64_t ip = off + (l == 0 ? N - 1 : l - 1);
    const double den = knot(f, off, N, T, l + 4) - knot(f, off, N, T, l);
#pragma unroll
    for (int d = 0; d < 3; d++) eco[d * f.Np + i] = 4.0 * (dco[d * f.Np + i] - dco[d * f.Np + ip]) / den;
}

// Filament of each node: the last filament whose first node is ≤ i.
__global__ void node_filament(const int64_t *__restrict__ off, int64_t Nf, int64_t Np, int32_t *__restrict__ fid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    int64_t lo = 0, hi = Nf - 1;  // off[0] = 0 ≤ i
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (off[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    fid[i] = (int32_t)lo;
}

struct Quad {
    int n;                      // quadrature points per segment (1 for straight segments)
    bool straight;              // NoQuadrature: midpoint of straight segments (pointdata_impl.jl:190-219)
    double gamma;               // lia_segment_fraction (1 if unset) for straight segments
    double z[16], w[16];        // rule on [0, 1]
    double lim_m[2], lim_p[2];  // integration limits of δ⁻ (segment i-1) and δ⁺ (segment i)
};

__global__ void point_data_kernel(Fil f, Quad qd, Box box, const double *__restrict__ X, const double *__restrict__ c,
                                  const double *__restrict__ dco, const double *__restrict__ eco,
                                  double *__restrict__ nodes, int32_t *__restrict__ idx_prev, double *__restrict__ sp,
                                  double *__restrict__ spp, double *__restrict__ dlm, double *__restrict__ dlp,
                                  double *__restrict__ pts, double *__restrict__ chg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t Np = f.Np;
    if (i >= Np) return;
    const int fi = f.fid[i];
    const int64_t off = f.off[fi];
    const int N = f.N[fi];
    const double T = f.T[fi];
    const int l = (int)(i - off);
    const int64_t ip = l == 0 ? off + N - 1 : i - 1;
    for (int d = 0; d < 3; d++) nodes[d * Np + i] = fold(X[d * Np + i], box.L[d]);
    idx_prev[i] = (int32_t)ip;
    double v[3];
    eval_spline<1>(f, fi, l, 0.0, dco, v);
    for (int d = 0; d < 3; d++) sp[d * Np + i] = v[d];
    eval_spline<2>(f, fi, l, 0.0, eco, v);
    for (int d = 0; d < 3; d++) spp[d * Np + i] = v[d];
    if (qd.straight) {
        // f[i-1], f[i+1] of the periodically extended filament
        const int64_t in = l == N - 1 ? off : i + 1;
        double xi[3], xp[3], xn[3];
        for (int d = 0; d < 3; d++) {
            const double xo = f.Xoff[d * f.Nf + fi];
            xi[d] = X[d * Np + i];
            xp[d] = X[d * Np + ip] - (l == 0 ? xo : 0.0);
            xn[d] = X[d * Np + in] + (l == N - 1 ? xo : 0.0);
        }
        double am = 0, ap = 0;
        for (int d = 0; d < 3; d++) {
            am += (xi[d] - xp[d]) * (xi[d] - xp[d]);
            ap += (xn[d] - xi[d]) * (xn[d] - xi[d]);
        }
        dlm[i] = qd.gamma * sqrt(am);
        dlp[i] = qd.gamma * sqrt(ap);
        for (int d = 0; d < 3; d++) {
            pts[d * Np + i] = fold((xi[d] + xn[d]) / 2, box.L[d]);
            chg[d * Np + i] = xn[d] - xi[d];
        }
        return;
    }
    // δ⁻, δ⁺ (integrate.jl:40-61 with Quadratures.integrate over (a, b))
    for (int side = 0; side < 2; side++) {
        const int seg = side == 0 ? (l == 0 ? N - 1 : l - 1) : l;
        const double a = side == 0 ? qd.lim_m[0] : qd.lim_p[0];
        const double b = side == 0 ? qd.lim_m[1] : qd.lim_p[1];
        const double delta = b - a;
        const double dt = knot(f, off, N, T, seg + 1) - knot(f, off, N, T, seg);
        double acc = 0.0;
        for (int q = 0; q < qd.n; q++) {
            eval_spline<1>(f, fi, seg, a + delta * qd.z[q], dco, v);
            acc = acc + qd.w[q] * sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
        (side == 0 ? dlm : dlp)[i] = dt * (delta * acc);
    }
    // quadrature points and charges of segment l
    const double dt = knot(f, off, N, T, l + 1) - knot(f, off, N, T, l);
    const int64_t Nq = (int64_t)qd.n * Np;
    for (int q = 0; q < qd.n; q++) {
        const int64_t j = (int64_t)qd.n * i + q;
        eval_spline<0>(f, fi, l, qd.z[q], c, v);
        for (int d = 0; d < 3; d++) pts[d * Nq + j] = fold(v[d], box.L[d]);
        eval_spline<1>(f, fi, l, qd.z[q], dco, v);
        const double wq = qd.w[q] * dt;
        for (int d = 0; d < 3; d++) chg[d * Nq + j] = wq * v[d];
    }
}

// Near-singular local integrals for lia_segment_fraction = γ
// (add_local_integrals!, local_integrals.jl:209-305): the parts of the two
// segments next to node i not covered by the LIA term, ζ ∈ [0, 1-γ] of segment
// i-1 and ζ ∈ [γ, 1] of segment i, integrated with the full Biot–Savart
// kernel and the near-singularity quadrature rule; FP64, added to the
// short-range outputs (v += Γ/4π Σ q × r/|r|³, ψ += Γ/4π Σ q/|r|).
struct LocalQuad {
    int n;
    double z[16], w[16];
    double lo[2], hi[2];
};

__global__ void local_integrals_kernel(Fil f, LocalQuad lq, Box box, const double *__restrict__ c,
                                       const double *__restrict__ dco, const double *__restrict__ nodes, double pref,
                                       double *__restrict__ v, double *__restrict__ psi) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t Np = f.Np;
    if (i >= Np) return;
    const int fi = f.fid[i];
    const int64_t off = f.off[fi];
    const int N = f.N[fi];
    const double T = f.T[fi];
    const int l = (int)(i - off);
    const double x[3] = {nodes[i], nodes[Np + i], nodes[2 * Np + i]};
    double av[3] = {0, 0, 0}, ap[3] = {0, 0, 0};
    for (int side = 0; side < 2; side++) {
        const int seg = side == 0 ? (l == 0 ? N - 1 : l - 1) : l;
        const double lo = lq.lo[side], delta = lq.hi[side] - lq.lo[side];
        const double dt = (knot(f, off, N, T, seg + 1) - knot(f, off, N, T, seg)) * delta;
        for (int q = 0; q < lq.n; q++) {
            const double y = lo + delta * lq.z[q];
            double sv[3], s1[3], r[3];
            eval_spline<0>(f, fi, seg, y, c, sv);
            eval_spline<1>(f, fi, seg, y, dco, s1);
            const double wq = lq.w[q] * dt;
            for (int d = 0; d < 3; d++) {
                const double sd = fold(sv[d], box.L[d]);
                double rd = x[d] - sd;
                if (rd >= box.Lh[d]) rd -= box.L[d];
                if (rd < -box.Lh[d]) rd += box.L[d];
                r[d] = rd;
                s1[d] *= wq;
            }
            const double rinv = 1.0 / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            const double rinv3 = rinv * rinv * rinv;
            av[0] += (s1[1] * r[2] - s1[2] * r[1]) * rinv3;
            av[1] += (s1[2] * r[0] - s1[0] * r[2]) * rinv3;
            av[2] += (s1[0] * r[1] - s1[1] * r[0]) * rinv3;
            for (int d = 0; d < 3; d++) ap[d] += s1[d] * rinv;
        }
    }
    for (int d = 0; d < 3; d++) {
        if (v) v[d * Np + i] += pref * av[d];
        if (psi) psi[d * Np + i] += pref * ap[d];
    }
}

// ----------------------------------------------------------------------------
// Time stepping on the device (Timestepping: SSPRK33.jl:17-39, advect.jl:20-72).

// Chordal knots of each filament: t_0 = 0, t_{i+1} = t_i + |X_{i+1} - X_i|, and the
// period T including the closing segment to X_0 + offset (one thread per filament).
__global__ void chordal_knots(Fil f, const double *__restrict__ X, double *__restrict__ ts, double *__restrict__ T) {
    const int64_t fi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (fi >= f.Nf) return;
    const int64_t off = f.off[fi];
    const int N = f.N[fi];
    const int64_t Np = f.Np;
    double acc = 0.0;
    ts[off] = 0.0;
    for (int i = 0; i < N; i++) {
        const int64_t a = off + i, b = (i == N - 1) ? off : a + 1;
        double d2 = 0.0;
        for (int d = 0; d < 3; d++) {
            const double xb = X[d * Np + b] + (i == N - 1 ? f.Xoff[d * f.Nf + fi] : 0.0);
            const double dd = xb - X[d * Np + a];
            d2 = d2 + dd * dd;
        }
        acc = acc + sqrt(d2);
        if (i < N - 1) ts[a + 1] = acc;
    }
    T[fi] = acc;
}

// Coefficients c of the periodic quintic interpolant of Y(t_i) = X_i - (t_i/T)·offset
// (spline.jl:163-193), one thread per filament. Row i of the collocation system
// holds B_{i-5+r,6}(t_i), r = 0..4, at columns (i-5+r) mod N; with unknowns
// shifted by 3 it is a cyclic 5-band matrix with the largest value on the
// diagonal. The band part is eliminated without pivoting (B-spline collocation
// matrices are totally positive) and the periodic corners are added back through
// the Woodbury identity with U = [e_0, e_1, e_{N-2}, e_{N-1}].
// buf: 12 arrays of Np doubles: 5 band diagonals, then 7 right-hand sides
// (3 components of Y, 4 columns of U).
__global__ void spline_solve(Fil f, const double *__restrict__ X, double *__restrict__ buf, double *__restrict__ c) {
    const int64_t fi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (fi >= f.Nf) return;
    const int64_t off = f.off[fi], Np = f.Np;
    const int N = f.N[fi];
    const double T = f.T[fi];
    double *band[5], *rhs[7];
    for (int r = 0; r < 5; r++) band[r] = buf + r * Np + off;
    for (int k = 0; k < 7; k++) rhs[k] = buf + (5 + k) * Np + off;
    double c0a = 0, c0b = 0, c1 = 0, c2 = 0, c3a = 0, c3b = 0;  // periodic corners
    for (int i = 0; i < N; i++) {
        const double ti = f.ts[off + i];
        double B[6];
        basis<6>(f, off, N, T, i, ti, B);
        for (int r = 0; r < 5; r++) band[r][i] = B[r];
        if (i == 0) {
            c0a = B[0], c0b = B[1];
            band[0][i] = band[1][i] = 0.0;
        }
        if (i == 1) {
            c1 = B[0];
            band[0][i] = 0.0;
        }
        if (i == N - 2) {
            c2 = B[4];
            band[4][i] = 0.0;
        }
        if (i == N - 1) {
            c3a = B[3], c3b = B[4];
            band[3][i] = band[4][i] = 0.0;
        }
        for (int d = 0; d < 3; d++) rhs[d][i] = X[d * Np + off + i] - (ti / T) * f.Xoff[d * f.Nf + fi];
        rhs[3][i] = i == 0 ? 1.0 : 0.0;
        rhs[4][i] = i == 1 ? 1.0 : 0.0;
        rhs[5][i] = i == N - 2 ? 1.0 : 0.0;
        rhs[6][i] = i == N - 1 ? 1.0 : 0.0;
    }
    // forward elimination (row i: band[r][i] at column i - 2 + r)
    for (int k = 0; k < N; k++) {
        const double piv = band[2][k];
        for (int sft = 1; sft <= 2 && k + sft < N; sft++) {
            const int i = k + sft;
            const double l = band[2 - sft][i] / piv;
            band[2 - sft][i] = 0.0;
            band[3 - sft][i] -= l * band[3][k];
            if (k + 2 < N) band[4 - sft][i] -= l * band[4][k];
            for (int q = 0; q < 7; q++) rhs[q][i] -= l * rhs[q][k];
        }
    }
    // back substitution, in place in rhs
    for (int k = N - 1; k >= 0; k--) {
        for (int q = 0; q < 7; q++) {
            double v = rhs[q][k];
            if (k + 1 < N) v -= band[3][k] * rhs[q][k + 1];
            if (k + 2 < N) v -= band[4][k] * rhs[q][k + 2];
            rhs[q][k] = v / band[2][k];
        }
    }
    // Woodbury: x = z - Z (I + Vᵀ Z)⁻¹ Vᵀ z, Vᵀ rows = the corner entries
    auto vt = [&](int a, const double *x) -> double {
        switch (a) {
            case 0: return c0a * x[N - 2] + c0b * x[N - 1];
            case 1: return c1 * x[N - 1];
            case 2: return c2 * x[0];
            default: return c3a * x[0] + c3b * x[1];
        }
    };
    double M[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) M[a][b] = (a == b ? 1.0 : 0.0) + vt(a, rhs[3 + b]);
    double g[4][3];
    for (int a = 0; a < 4; a++)
        for (int d = 0; d < 3; d++) g[a][d] = vt(a, rhs[d]);
    // solve M w = g (4×4, partial pivoting)
    int perm[4] = {0, 1, 2, 3};
    for (int k = 0; k < 4; k++) {
        int p = k;
        for (int r = k + 1; r < 4; r++)
            if (fabs(M[perm[r]][k]) > fabs(M[perm[p]][k])) p = r;
        const int tmp = perm[k];
        perm[k] = perm[p];
        perm[p] = tmp;
        for (int r = k + 1; r < 4; r++) {
            const double l = M[perm[r]][k] / M[perm[k]][k];
            for (int b = k; b < 4; b++) M[perm[r]][b] -= l * M[perm[k]][b];
            for (int d = 0; d < 3; d++) g[perm[r]][d] -= l * g[perm[k]][d];
        }
    }
    double w[4][3];
    for (int k = 3; k >= 0; k--)
        for (int d = 0; d < 3; d++) {
            double v = g[perm[k]][d];
            for (int b = k + 1; b < 4; b++) v -= M[perm[k]][b] * w[b][d];
            w[k][d] = v / M[perm[k]][k];
        }
    // unknown m = (j + 3) mod N holds coefficient j
    for (int j = 0; j < N; j++) {
        int m = j + 3;
        if (m >= N) m -= N;
        for (int d = 0; d < 3; d++) {
            double v = rhs[d][m];
            for (int b = 0; b < 4; b++) v -= rhs[3 + b][m] * w[b][d];
            c[d * Np + off + j] = v;
        }
    }
}

// out = x0 + a·v over 3·Np values
__global__ void axpy_nodes(int64_t n, const double *__restrict__ x0, double a, const double *__restrict__ v,
                           double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = x0[i] + a * v[i];
}

// SSPRK33 velocity combinations: mode 0: vs = v1 + v2; mode 1: vs = (vs + 4·v3)/6
__global__ void rk_combine(int64_t n, int mode, double *__restrict__ vs, const double *__restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vs[i] = mode == 0 ? vs[i] + v[i] : (vs[i] + 4.0 * v[i]) / 6.0;
}

}  // namespace

Fil filament_view(const vpbs_ctx *c) {
    Fil f;
    f.fid = c->fil_fid;
    f.off = c->fil_off;
    f.N = c->fil_n;
    f.T = c->fil_T;
    f.Xoff = c->fil_xoff;
    f.ts = c->fil_ts;
    f.Np = c->Np;
    f.Nf = c->pd_nf;
    return f;
}

// Point data from the filament state on the device (nodes, knots, periods,
// offsets, spline coefficients), with the quadrature set by vpbs_set_filaments.
int point_data_from_state(vpbs_ctx *c) {
    hipStream_t s = c->s_sr;
    const Fil f = filament_view(c);
    const int64_t Np = c->Np;
    Quad qd{};
    qd.n = c->p.nquad;
    qd.straight = c->pd_straight;
    qd.gamma = std::isnan(c->pd_gamma) ? 1.0 : c->pd_gamma;
    for (int q = 0; q < qd.n && !qd.straight; q++) {
        qd.z[q] = c->pd_z[q];
        qd.w[q] = c->pd_w[q];
    }
    if (std::isnan(c->pd_gamma)) {
        qd.lim_m[0] = 0.0;
        qd.lim_m[1] = 1.0;
        qd.lim_p[0] = 0.0;
        qd.lim_p[1] = 1.0;
    } else {  // lia_integration_limits (local_integrals.jl:100-101)
        const double g = c->pd_gamma;
        qd.lim_m[0] = 1.0 - g;
        qd.lim_m[1] = 1.0;
        qd.lim_p[0] = 0.0;
        qd.lim_p[1] = g;
    }
    c->have_points = false;
    stage_begin(c, ST_POINTS, s);
    const unsigned nb = nblocks(Np, 256);
    derivative_coefs<<<nb, 256, 0, s>>>(f, c->fil_c, c->fil_d);
    second_derivative_coefs<<<nb, 256, 0, s>>>(f, c->fil_d, c->fil_e);
    point_data_kernel<<<nblocks(Np, 128), 128, 0, s>>>(f, qd, c->box, c->fil_x, c->fil_c, c->fil_d, c->fil_e,
                                                       c->nodes, c->idx_prev, c->sp, c->spp, c->dlm, c->dlp,
                                                       c->points, c->charges);
    VPBS_HIP(c, hipGetLastError());
    stage_end(c, ST_POINTS, s);
    VPBS_HIP(c, hipStreamSynchronize(s));
    VPBS_TRY(device_check(c));
    float ms = 0.f;
    VPBS_HIP(c, hipEventElapsedTime(&ms, c->ev_beg[ST_POINTS], c->ev_end[ST_POINTS]));
    c->timings.point_data = ms;
    c->have_points = true;
    c->points_version++;
    c->have_stats = false;
    c->last_vel = c->last_psi = false;
    return VPBS_OK;
}

// Upload the filament description, then derive the point data on the device.
// Argument meaning as in vpbs_set_filaments (include/vpbs.h); the host checks
// done here make every index the kernels form stay inside its array.
int build_point_data(vpbs_ctx *c, int64_t Nf, const int64_t *counts, const double *X, const double *knots,
                     const double *periods, const double *offsets, const double *coefs, const double *quad_nodes,
                     const double *quad_weights, double lia_segment_fraction, int32_t nquad_ns,
                     const double *quad_ns_nodes, const double *quad_ns_weights) {
    if (Nf < 1) return fail(c, VPBS_ERR_INVALID, "need at least one filament (got Nf = %lld)", (long long)Nf);
    if (!counts || !X || !knots || !periods || !offsets || !coefs)
        return fail(c, VPBS_ERR_INVALID, "null input array");
    if ((quad_nodes == nullptr) != (quad_weights == nullptr))
        return fail(c, VPBS_ERR_INVALID, "quad_nodes and quad_weights must both be given or both be NULL");
    const bool straight = quad_nodes == nullptr;
    const int nq = c->p.nquad;
    if (straight && nq != 1)
        return fail(c, VPBS_ERR_INVALID, "straight segments (no quadrature rule) need nquad = 1 (got %d)", nq);
    if (!std::isnan(lia_segment_fraction) && !(lia_segment_fraction > 0.0 && lia_segment_fraction <= 1.0))
        return fail(c, VPBS_ERR_INVALID, "lia_segment_fraction must be in (0, 1] (got %g)", lia_segment_fraction);
    if (nquad_ns < 0 || nquad_ns > 16 || (nquad_ns > 0 && (!quad_ns_nodes || !quad_ns_weights)))
        return fail(c, VPBS_ERR_INVALID, "near-singularity quadrature: 0..16 points with nodes and weights");
    std::vector<int64_t> off((size_t)Nf);
    std::vector<int32_t> nn((size_t)Nf);
    int64_t Np = 0;
    for (int64_t f = 0; f < Nf; f++) {
        if (counts[f] < 5 || counts[f] > INT32_MAX)
            return fail(c, VPBS_ERR_INVALID, "filament %lld has %lld nodes; quintic splines need at least 5",
                        (long long)f, (long long)counts[f]);
        off[(size_t)f] = Np;
        nn[(size_t)f] = (int32_t)counts[f];
        Np += counts[f];
        if (Np > INT32_MAX) return fail(c, VPBS_ERR_INVALID, "too many nodes for 32-bit indices");
        if (!(periods[f] > 0.0) || !std::isfinite(periods[f]))
            return fail(c, VPBS_ERR_INVALID, "period of filament %lld is %g (must be finite and > 0)",
                        (long long)f, periods[f]);
    }
    const int64_t Nq = (int64_t)nq * Np;
    if (Nq > INT32_MAX / 2) return fail(c, VPBS_ERR_INVALID, "too many quadrature points for 32-bit indices");
    VPBS_HIP(c, hipSetDevice(c->device));
    c->have_points = false;
    VPBS_TRY(ensure_capacity(c, Np, Nq));
    if (Np > c->cap_fil_np) {
        double **bufs[] = {&c->fil_x, &c->fil_c, &c->fil_d, &c->fil_e};
        for (double **b : bufs) {
            if (*b) VPBS_HIP(c, hipFree(*b));
            *b = nullptr;
            VPBS_HIP(c, hipMalloc((void **)b, sizeof(double) * 3 * (size_t)Np));
        }
        if (c->fil_ts) VPBS_HIP(c, hipFree(c->fil_ts));
        if (c->fil_fid) VPBS_HIP(c, hipFree(c->fil_fid));
        c->fil_ts = nullptr;
        c->fil_fid = nullptr;
        VPBS_HIP(c, hipMalloc((void **)&c->fil_ts, sizeof(double) * (size_t)Np));
        VPBS_HIP(c, hipMalloc((void **)&c->fil_fid, sizeof(int32_t) * (size_t)Np));
        c->cap_fil_np = Np;
    }
    if (Nf > c->cap_fil_nf) {
        if (c->fil_off) VPBS_HIP(c, hipFree(c->fil_off));
        if (c->fil_n) VPBS_HIP(c, hipFree(c->fil_n));
        if (c->fil_T) VPBS_HIP(c, hipFree(c->fil_T));
        if (c->fil_xoff) VPBS_HIP(c, hipFree(c->fil_xoff));
        c->fil_off = nullptr;
        c->fil_n = nullptr;
        c->fil_T = c->fil_xoff = nullptr;
        VPBS_HIP(c, hipMalloc((void **)&c->fil_off, sizeof(int64_t) * (size_t)Nf));
        VPBS_HIP(c, hipMalloc((void **)&c->fil_n, sizeof(int32_t) * (size_t)Nf));
        VPBS_HIP(c, hipMalloc((void **)&c->fil_T, sizeof(double) * (size_t)Nf));
        VPBS_HIP(c, hipMalloc((void **)&c->fil_xoff, sizeof(double) * 3 * (size_t)Nf));
        c->cap_fil_nf = Nf;
    }
    hipStream_t s = c->s_sr;
    const size_t d3 = sizeof(double) * 3 * (size_t)Np;
    VPBS_HIP(c, hipMemcpyAsync(c->fil_x, X, d3, hipMemcpyHostToDevice, s));
    VPBS_HIP(c, hipMemcpyAsync(c->fil_c, coefs, d3, hipMemcpyHostToDevice, s));
    VPBS_HIP(c, hipMemcpyAsync(c->fil_ts, knots, sizeof(double) * (size_t)Np, hipMemcpyHostToDevice, s));
    VPBS_HIP(c, hipMemcpyAsync(c->fil_off, off.data(), sizeof(int64_t) * (size_t)Nf, hipMemcpyHostToDevice, s));
    VPBS_HIP(c, hipMemcpyAsync(c->fil_n, nn.data(), sizeof(int32_t) * (size_t)Nf, hipMemcpyHostToDevice, s));
    VPBS_HIP(c, hipMemcpyAsync(c->fil_T, periods, sizeof(double) * (size_t)Nf, hipMemcpyHostToDevice, s));
    VPBS_HIP(c, hipMemcpyAsync(c->fil_xoff, offsets, sizeof(double) * 3 * (size_t)Nf, hipMemcpyHostToDevice, s));

    // quadrature and local-integral configuration, kept for rebuilds from the device state
    c->pd_nf = Nf;
    c->pd_straight = straight;
    c->pd_gamma = lia_segment_fraction;
    for (int q = 0; q < nq && !straight; q++) {
        c->pd_z[q] = quad_nodes[q];
        c->pd_w[q] = quad_weights[q];
    }
    c->local_ns = (!std::isnan(lia_segment_fraction) && nquad_ns > 0) ? nquad_ns : 0;
    c->local_gamma = lia_segment_fraction;
    c->local_nf = Nf;
    for (int q = 0; q < c->local_ns; q++) {
        c->local_z[q] = quad_ns_nodes[q];
        c->local_w[q] = quad_ns_weights[q];
    }
    node_filament<<<nblocks(Np, 256), 256, 0, s>>>(c->fil_off, Nf, Np, c->fil_fid);
    VPBS_HIP(c, hipGetLastError());
    c->Np = Np;
    c->Nq = Nq;
    c->have_filaments = true;
    return point_data_from_state(c);
}

int add_local_integrals(vpbs_ctx *c, uint32_t flags, hipStream_t s) {
    if (!c->local_ns) return VPBS_OK;
    const bool vel = flags & VPBS_VELOCITY, psi = flags & VPBS_STREAMFUNCTION;
    const Fil f = filament_view(c);
    LocalQuad lq{};
    lq.n = c->local_ns;
    for (int q = 0; q < lq.n; q++) {
        lq.z[q] = c->local_z[q];
        lq.w[q] = c->local_w[q];
    }
    lq.lo[0] = 0.0;  // nonlia_integration_limits (local_integrals.jl:103-104)
    lq.hi[0] = 1.0 - c->local_gamma;
    lq.lo[1] = c->local_gamma;
    lq.hi[1] = 1.0;
    const double pref = c->p.Gamma / (4 * M_PI);
    local_integrals_kernel<<<nblocks(c->Np, 128), 128, 0, s>>>(f, lq, c->box, c->fil_c, c->fil_d, c->nodes, pref,
                                                              vel ? c->v_sr : nullptr, psi ? c->psi_sr : nullptr);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

// SSPRK33 steps of the Biot–Savart velocity on the device-resident filaments
// (Timestepping.jl:1038-1076 without forcing, reconnections or adaptivity).
int advance_filaments(vpbs_ctx *c, double dt, int nsteps) {
    const int64_t Np = c->Np, n3 = 3 * Np;
    if (n3 > c->cap_step) {
        if (c->step_buf) VPBS_HIP(c, hipFree(c->step_buf));
        c->step_buf = nullptr;
        VPBS_HIP(c, hipMalloc((void **)&c->step_buf, sizeof(double) * 3 * (size_t)n3));
        c->cap_step = n3;
    }
    if (12 * Np > c->cap_solve) {
        if (c->solve_buf) VPBS_HIP(c, hipFree(c->solve_buf));
        c->solve_buf = nullptr;
        VPBS_HIP(c, hipMalloc((void **)&c->solve_buf, sizeof(double) * 12 * (size_t)Np));
        c->cap_solve = 12 * Np;
    }
    double *X0 = c->step_buf, *V1 = X0 + n3, *Vs = V1 + n3;
    hipStream_t s = c->s_sr;
    const Fil f = filament_view(c);
    const unsigned nbn = nblocks(n3, 256), nbf = nblocks(c->pd_nf, 64);
    // node velocities into c->v_out from the current state: the learned evaluation (cheap
    // splitting, correction, sampled check) when a network is set up, else the plain one
    const bool learned = c->lc_ready && c->lc_net;
    auto velocity = [&]() -> int {
        VPBS_TRY(point_data_from_state(c));
        if (learned) return vpbs_compute_learned(c, nullptr, nullptr);
        return vpbs_compute(c, VPBS_VELOCITY, __builtin_nan(""), nullptr, nullptr);
    };
    auto stage = [&](double a, const double *v) -> int {  // X = X0 + a·v, coefficients with the same knots
        axpy_nodes<<<nbn, 256, 0, s>>>(n3, X0, a, v, c->fil_x);
        spline_solve<<<nbf, 64, 0, s>>>(f, c->fil_x, c->solve_buf, c->fil_c);
        VPBS_HIP(c, hipGetLastError());
        return VPBS_OK;
    };
    for (int step = 0; step < nsteps; step++) {
        VPBS_TRY(velocity());
        VPBS_HIP(c, hipMemcpyAsync(X0, c->fil_x, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
        VPBS_HIP(c, hipMemcpyAsync(V1, c->v_out, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
        VPBS_TRY(stage(dt, V1));
        VPBS_TRY(velocity());
        VPBS_HIP(c, hipMemcpyAsync(Vs, V1, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
        rk_combine<<<nbn, 256, 0, s>>>(n3, 0, Vs, c->v_out);
        VPBS_TRY(stage(dt / 4, Vs));
        VPBS_TRY(velocity());
        rk_combine<<<nbn, 256, 0, s>>>(n3, 1, Vs, c->v_out);
        // final advection: new knots (chordal), then coefficients
        axpy_nodes<<<nbn, 256, 0, s>>>(n3, X0, dt, Vs, c->fil_x);
        chordal_knots<<<nbf, 64, 0, s>>>(f, c->fil_x, c->fil_ts, c->fil_T);
        spline_solve<<<nbf, 64, 0, s>>>(f, c->fil_x, c->solve_buf, c->fil_c);
        VPBS_HIP(c, hipGetLastError());
    }
    return point_data_from_state(c);  // point data of the final state
}

}  // namespace vpbs
