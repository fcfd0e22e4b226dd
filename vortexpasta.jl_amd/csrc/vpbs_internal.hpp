// Internal declarations of libvpbs (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/vpbs.h"

namespace vpbs {

// Geometry of the periodic box.
struct Box {
    double L[3];
    double Lh[3];
};

// Deterministic sort scratch (grow-only device buffers).
struct SortBuffers {
    int32_t *rank = nullptr;   // [n] rank within bin from the histogram pass
    int64_t cap_n = 0;
    int32_t *scan_tmp = nullptr;  // block sums for the multi-level scan
    int64_t cap_tmp = 0;
    // radix sort: sorted keys, then the original indices ([2n]); library scratch (bytes)
    int32_t *keys_out = nullptr;
    int64_t cap_radix = 0;
    unsigned char *radix_tmp = nullptr;
    int64_t cap_radix_tmp = 0;
};

// Slabs of the interpolation fused with the inverse x transform: SLAB_HY y-rows ×
// SLAB_HZ z-rows × all x-cells (spectral.hip; nodes sorted by slab in longrange.hip).
constexpr int SLAB_HY = 4, SLAB_HZ = 2;
// Node record of that interpolation: the W kernel weights per direction, padded to
// whole 16-byte vectors (elem = bytes per value).
constexpr int slab_rec_len(int W, int elem) { return elem == 4 ? (3 * W + 3) & ~3 : (3 * W + 1) & ~1; }

// Per-stage event pairs.
enum Stage {
    ST_LOCAL = 0,
    ST_CELL,
    ST_PAIR,
    ST_SELF,
    ST_LRSORT,
    ST_SPREAD,
    ST_FFTF,
    ST_FOURIER,
    ST_FFTB,
    ST_INTERP,
    ST_MERGE,
    ST_NDSORT,   // slab sort and records of the nodes (short-range stream, read by the interpolation)
    ST_FEATURES, // learned correction: node cell list and feature kernel
    ST_MLP,      // learned correction: network inference and the corrected sum
    ST_CHECK,    // learned correction: sampled check against the converged setting
    ST_POINTS,   // vpbs_set_filaments (not part of vpbs_compute)
    ST_COUNT
};

}  // namespace vpbs

struct vpbs_ctx {
    vpbs_params p{};
    std::string err;
    int device = 0;
    hipStream_t s_sr = nullptr, s_lr = nullptr;
    hipEvent_t ev_beg[vpbs::ST_COUNT]{}, ev_end[vpbs::ST_COUNT]{};
    bool stage_ran[vpbs::ST_COUNT]{};
    hipEvent_t ev_lr_done = nullptr;
    hipEvent_t ev_nodes = nullptr;  // slab-sorted node records ready (recorded on s_sr, waited on by s_lr)
    vpbs_timings timings{};
    vpbs::Box box{};

    // point data (device, FP64)
    int64_t Np = 0, Nq = 0, cap_np = 0, cap_nq = 0;
    bool have_points = false;
    double *nodes = nullptr, *sp = nullptr, *spp = nullptr, *dlm = nullptr, *dlp = nullptr;
    int32_t *idx_prev = nullptr;
    double *points = nullptr, *charges = nullptr;

    // filament splines (device, vpbs_set_filaments)
    int64_t cap_fil_np = 0, cap_fil_nf = 0;
    double *fil_x = nullptr, *fil_ts = nullptr, *fil_c = nullptr, *fil_d = nullptr, *fil_e = nullptr;
    int32_t *fil_fid = nullptr, *fil_n = nullptr;
    int64_t *fil_off = nullptr;
    double *fil_T = nullptr, *fil_xoff = nullptr;
    bool have_filaments = false;   // filament state on the device (vpbs_set_filaments)
    int64_t pd_nf = 0;             // filaments of that state
    bool pd_straight = false;      // NoQuadrature
    double pd_gamma = 0;           // lia_segment_fraction (NaN: none)
    double pd_z[16]{}, pd_w[16]{}; // quadrature rule on [0, 1]
    double *step_buf = nullptr;    // time stepping: X0, V1, Vs (3·Np each)
    int64_t cap_step = 0;
    double *solve_buf = nullptr;   // spline solve: 5 band + 7 right-hand sides per node
    int64_t cap_solve = 0;
    // near-singular local integrals (lia_segment_fraction), set by vpbs_set_filaments
    int local_ns = 0;          // near-singularity quadrature points (0: not evaluated on the device)
    double local_gamma = 0;
    int64_t local_nf = 0;
    double local_z[16]{}, local_w[16]{};

    // outputs (device, FP64, 3*Np each)
    double *v_sr = nullptr, *psi_sr = nullptr, *v_lr = nullptr, *psi_lr = nullptr, *v_out = nullptr,
           *psi_out = nullptr;
    bool last_vel = false, last_psi = false;

    // cell lists
    int32_t Nc[3] = {0, 0, 0};
    int64_t ncells = 0;
    int32_t *src_key = nullptr, *src_start = nullptr, *src_perm = nullptr;
    int32_t *tgt_key = nullptr, *tgt_start = nullptr, *tgt_perm = nullptr;
    int32_t *tgt_cell = nullptr;  // per node (original order): cell (ix, iy, iz, key), 4*Np
    double *src_pos = nullptr;   // sorted sources, 3*Nq (component-major)
    void *src_chg = nullptr;     // sorted charges, 3*Nq of the compute type
    int32_t *src_idx = nullptr;  // original index of each sorted source
    int64_t *stat_counts = nullptr;
    int32_t *stat_cands = nullptr;  // per node: candidate sources tested by the pair kernel (stats calls)
    uint64_t *stat_hashes = nullptr;
    bool have_stats = false;
    vpbs::SortBuffers sort_sr, sort_tg, sort_lr, sort_nd;

    // long range
    int32_t G[3] = {0, 0, 0};   // oversampled grid
    int64_t comp_stride = 0;    // real grid values per component (G0·G1·G2)
    int32_t ntile[3] = {0, 0, 0};
    int64_t ntiles = 0;
    // interpolation fused with the inverse x transform (VPBS_FFT_FUSED): the grid is cut
    // into slabs of SLAB_HY y-rows × SLAB_HZ z-rows × all x-cells; nodes are sorted by
    // the slab of their first stencil cell (nd_*), each slab writes its part of the
    // stencil sums of the nodes it reaches into slab_part ([slots][3][Np])
    bool slab_interp = false;
    int64_t nslabs = 0;
    int64_t cap_nd = 0;
    int32_t *nd_key = nullptr, *nd_start = nullptr, *nd_perm = nullptr, *nd_i0 = nullptr;
    void *nd_rec = nullptr, *slab_part = nullptr;
    // Real grid, 3 components, rows of y = 2p and y = 2p+1 interleaved as one
    // complex row (pair p): value (x, y, z) at ((z·G1/2 + y/2)·G0 + x)·2 + (y & 1).
    // Spreading writes it, the inverse x transform rewrites it, interpolation reads it.
    void *grid = nullptr;
    // Pencils of retained modes (complex), same sizes in both transform paths:
    //   fused:  ypen = x1 [3][G2][G1][nkx], zpen = x2 [3][nky][G2][nkx] (z space on both
    //           sides of the z transforms: vorticity in, velocity out);
    //   rocFFT: ypen [3][nkx][G2][G1], zpen [3][nkx][nky][G2] (vorticity, then velocity, in place).
    // The streamfunction gets its own zpen2 (allocated on demand).
    int32_t nkx = 0, nky = 0;
    int64_t ypen_stride = 0, zpen_stride = 0;  // complex values per component
    void *ypen = nullptr, *zpen = nullptr, *zpen2 = nullptr;
    // unsmoothed field kept by VPBS_KEEP_UHAT: [3][nky][nkz][nkx] complex (compute type),
    // ky = jy (jy ≤ K1) or jy - nky, kz likewise with K2 and nkz
    int32_t nkz = 0;
    int64_t uhat_stride = 0;   // complex values per component
    void *uhat = nullptr;
    int uhat_quantity = 0;     // 0: none, VPBS_UHAT_VELOCITY or VPBS_UHAT_STREAMFUNCTION
    double *spec_scratch = nullptr;  // this GPU's `scratch` array (allocation-free spectrum)
    bool pair_tiled = false;   // tiled pair kernel (VPBS_PAIR=tile at creation; default one wavefront per target)
    int32_t *lr_key = nullptr, *lr_start = nullptr, *lr_perm = nullptr;
    uint8_t *lr_active = nullptr;  // per tile: 1 if any neighbour tile holds a point
    int32_t *lr_i0 = nullptr;   // tile-sorted points: first stencil index per direction, 3*Nq
    void *lr_w = nullptr;       // tile-sorted points: one record per point (3·2m weights, charge; compute type)
    double *deconv = nullptr;   // 1/ψ̂ tables for retained kx, ky, kz (concatenated)
    int32_t kmax_mode[3] = {0, 0, 0};
    double es_beta = 0;
    // rocFFT: batched contiguous 1D c2c transforms along x, y and z, forward and inverse
    rocfft_plan plan[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    rocfft_execution_info fft_info = nullptr;
    void *fft_work = nullptr;
    size_t fft_work_size = 0;
    bool plans_ready = false;
    // fused transforms (spectral.hip): twiddle tables e^{-2πin/G_d} (compute type) and the
    // separable factors (1/ψ̂)² e^{-k_d²/4α²} of the Fourier-space kernel (FP64)
    bool fft_fused = false;
    size_t spec_lds_max = 0;    // LDS per block available to the fused kernels
    void *spec_tw = nullptr;
    double *spec_fac = nullptr;

    // point-data generation (incremented whenever the point data changes) and the generation
    // the cell lists and the kept Fourier field were last computed for
    uint64_t points_version = 0, cells_version = ~0ull, uhat_version = ~0ull;

    // learned correction (learned.hip)
    vpbs_mlp *lc_net = nullptr;       // borrowed from the caller
    bool lc_ready = false, lc_active = false;
    double lc_alpha_c = 0, lc_alpha_f = 0, lc_rcut_f = 0, lc_rfeat = 0, lc_frac = 0, lc_tol = 0;
    uint64_t lc_seed = 0, lc_calls = 0;
    int64_t lc_fallbacks = 0;         // rejected corrections since vpbs_learned_setup
    double lc_max_err = 0.0;          // largest sampled error of the corrected velocity since then
    int64_t lc_cap = 0, lc_cap_s = 0;
    double *lc_npos = nullptr;        // cell-sorted node positions [3][Np]
    float *lc_rec = nullptr;          // cell-sorted strand records [Np][12] (tangent, index, neighbour offsets)
    void *lc_nset = nullptr;          // cell-sorted node set-up records [Np] × 128 B (frame, filament, cell)
    float *lc_feat = nullptr;         // features [Np][128]
    float *lc_frame = nullptr;        // local frames (t, n, b) [Np][9]
    float *lc_y = nullptr;            // network outputs [Np][3]
    double *lc_dvp = nullptr;         // applied correction [3][Np]
    double *lc_dsr = nullptr;         // converged − cheap short-range difference [3][n] (sample or all nodes)
    double *lc_dlr = nullptr;         // converged − cheap long-range difference [3][n]
    int32_t *lc_sidx = nullptr;       // sample nodes
    double *lc_spos = nullptr;        // their positions [3][ns]
    double *lc_stats = nullptr;       // check sums (3 doubles, 1 spare), then per-block Σ|v|² partials
    hipEvent_t lc_ev = nullptr;
    hipEvent_t lc_ev_cells = nullptr;  // short-range cell lists built (learned evaluation)
    int64_t lc_ns = 0;                // sample size of the current learned evaluation
};

namespace vpbs {

int fail(vpbs_ctx *c, int code, const char *fmt, ...);

#define VPBS_HIP(ctx, expr)                                                                        \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::vpbs::fail((ctx), VPBS_ERR_HIP, "%s failed: %s (%s:%d)", #expr,               \
                                hipGetErrorString(e_), __FILE__, __LINE__);                        \
    } while (0)

#define VPBS_TRY(expr)          \
    do {                        \
        int r_ = (expr);        \
        if (r_ != VPBS_OK)      \
            return r_;          \
    } while (0)

template <typename T>
int grow(vpbs_ctx *c, T **ptr, int64_t *cap, int64_t n) {
    if (n <= *cap && *ptr) return VPBS_OK;
    if (*ptr) {
        VPBS_HIP(c, hipFree(*ptr));
        *ptr = nullptr;
    }
    int64_t m = n > 0 ? n : 1;
    VPBS_HIP(c, hipMalloc((void **)ptr, sizeof(T) * (size_t)m));
    *cap = m;
    return VPBS_OK;
}

// sort.hip
// Order by key; with within_bin = true also by original index inside each bin (deterministic),
// otherwise the order inside a bin is that of the atomic ranks (for callers to which it does
// not matter). start[b] = first position of bin b, start[nbins] = n.
int counting_sort(vpbs_ctx *c, SortBuffers &sb, const int32_t *keys, int64_t n, int64_t nbins, int32_t *start,
                  int32_t *perm, hipStream_t s, bool within_bin = true);
// The same result (order by key, then original index; start[b] = first position of
// bin b, start[nbins] = n) by a stable LSD radix sort: no per-bin pass, for large bins.
int radix_sort_bins(vpbs_ctx *c, SortBuffers &sb, const int32_t *keys, int64_t n, int64_t nbins, int32_t *start,
                    int32_t *perm, hipStream_t s);

// shortrange.hip
int shortrange_compute(vpbs_ctx *c, uint32_t flags, double integration_cutoff);

// longrange.hip
int longrange_setup(vpbs_ctx *c);
int longrange_compute(vpbs_ctx *c, uint32_t flags);
int longrange_get_uhat(vpbs_ctx *c, int32_t *quantity, int64_t *dims3, double *out);

// fused spectral pipeline (spectral.hip): grid → x1 → x2 (forward x, y), the z
// transforms with the Fourier-space kernel on x2 (velocity) / zpen2 (streamfunction),
// and x2 → x1 → grid (inverse y, x); x1 = ypen, x2 = zpen.
bool spectral_supported(const vpbs_ctx *c);
int spectral_setup(vpbs_ctx *c);
int spectral_forward(vpbs_ctx *c);
int spectral_fourier(vpbs_ctx *c, uint32_t flags);
int spectral_inverse(vpbs_ctx *c, void *zp);
// The inverse y transform alone (zp → x1), and the inverse x transform fused with the
// interpolation of x1 onto the nodes (out: FP64 [3][Np]; slab-sorted node data nd_*).
bool spectral_slab_supported(const vpbs_ctx *c);
int spectral_slab_slots(const vpbs_ctx *c);
int spectral_y_inverse(vpbs_ctx *c, void *zp);
int spectral_x_inverse_interp(vpbs_ctx *c, double *out);
int spectral_band(vpbs_ctx *c, double kmin2, double kmax2, double cf = 0, double cc = 0);
void spectral_release(vpbs_ctx *c);
int longrange_spectrum(vpbs_ctx *c, int64_t nk, double dk, double *Ek);
int longrange_spectrum_noalloc(vpbs_ctx *c, int64_t nk, double dk, double *Ek);
int longrange_band(vpbs_ctx *c, double kmin, double kmax, double *v_out);
// The kept field filtered by (e^{−cf·k²} − e^{−cc·k²}) on every retained mode, transformed
// to the real grid (long-range stream); then interpolated at n positions pos[3][n] (device,
// folded) into out[3][n] (device, FP64).
int longrange_filtered_grid(vpbs_ctx *c, double cf, double cc);
int longrange_interp_at(vpbs_ctx *c, int64_t n, const double *pos, double *out);
void longrange_release(vpbs_ctx *c);
int oversampled_size(int N, double sigma);
double es_beta(int m, double sigma);

// pointdata.hip
int build_point_data(vpbs_ctx *c, int64_t Nf, const int64_t *counts, const double *X, const double *knots,
                     const double *periods, const double *offsets, const double *coefs, const double *quad_nodes,
                     const double *quad_weights, double lia_segment_fraction, int32_t nquad_ns,
                     const double *quad_ns_nodes, const double *quad_ns_weights);
int add_local_integrals(vpbs_ctx *c, uint32_t flags, hipStream_t s);
int point_data_from_state(vpbs_ctx *c);
int advance_filaments(vpbs_ctx *c, double dt, int nsteps);

// vpbs_api.hip
int ensure_capacity(vpbs_ctx *c, int64_t Np, int64_t Nq);
int device_check(vpbs_ctx *c);

// shortrange.hip: target (node) cell list of the current point data (tgt_key/start/perm/cell)
int build_target_cells(vpbs_ctx *c, hipStream_t s);

// mlp.hip
int mlp_infer(vpbs_mlp *m, int64_t n, const float *X, float *Y, hipStream_t s);
const char *mlp_error(const vpbs_mlp *m);
int mlp_n_in(const vpbs_mlp *m);
int mlp_n_out(const vpbs_mlp *m);

// learned.hip: stages of the learned evaluation inside vpbs_compute (short-range stream)
int learned_after_shortrange(vpbs_ctx *c);
int learned_after_merge(vpbs_ctx *c);
int learned_check_long(vpbs_ctx *c);
int learned_check_short(vpbs_ctx *c);
void learned_release(vpbs_ctx *c);

// helpers
void stage_begin(vpbs_ctx *c, Stage st, hipStream_t s);
void stage_end(vpbs_ctx *c, Stage st, hipStream_t s);
void collect_timings(vpbs_ctx *c);
// Synchronises both streams of the context when it goes out of scope.
struct StreamDrain {
    vpbs_ctx *c;
    ~StreamDrain();
};

}  // namespace vpbs
