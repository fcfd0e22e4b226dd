// C ABI of libvpbs: context life cycle, point upload and the compute entry points.
//
// vpbs_compute is the device part of _compute_on_nodes!
// (src/BiotSavart/BiotSavart.jl:515-612): the short-range task (LIA, pairs,
// self-interaction; do_shortrange!, BiotSavart.jl:468-513) runs on one HIP
// stream and the long-range task (do_longrange!, BiotSavart.jl:406-466) on a
// second one, replacing the reference's two Julia tasks; the results are
// summed on the device.
#include "vpbs_internal.hpp"

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

namespace vpbs {

static std::string g_create_error;

int fail(vpbs_ctx *c, int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c)
        c->err = buf;
    else
        g_create_error = buf;
    return code;
}

void stage_begin(vpbs_ctx *c, Stage st, hipStream_t s) {
    (void)hipEventRecord(c->ev_beg[st], s);
    c->stage_ran[st] = true;
}

void stage_end(vpbs_ctx *c, Stage st, hipStream_t s) { (void)hipEventRecord(c->ev_end[st], s); }

int merge_outputs(vpbs_ctx *c, const double *a, const double *b, double *o, int64_t n, hipStream_t s);

StreamDrain::~StreamDrain() {
    if (c->s_sr) (void)hipStreamSynchronize(c->s_sr);
    if (c->s_lr) (void)hipStreamSynchronize(c->s_lr);
}

// Stage times of the last compute (ms; 0 for stages that did not run).
void collect_timings(vpbs_ctx *c) {
    vpbs_timings &T = c->timings;
    struct {
        Stage st;
        double *slot;
    } tab[] = {{ST_LOCAL, &T.local_term}, {ST_CELL, &T.cell_build}, {ST_PAIR, &T.pair},
               {ST_SELF, &T.self},        {ST_LRSORT, &T.lr_sort},   {ST_SPREAD, &T.spread},
               {ST_FFTF, &T.fft_forward}, {ST_FOURIER, &T.fourier},  {ST_FFTB, &T.fft_backward},
               {ST_INTERP, &T.interp},    {ST_MERGE, &T.merge},      {ST_NDSORT, &T.node_sort},
               {ST_FEATURES, &T.features}, {ST_MLP, &T.mlp},         {ST_CHECK, &T.check}};
    for (auto &e : tab) {
        float ms = 0.f;
        *e.slot = 0.0;
        if (c->stage_ran[e.st] && hipEventElapsedTime(&ms, c->ev_beg[e.st], c->ev_end[e.st]) == hipSuccess)
            *e.slot = ms;
    }
}

static int validate(const vpbs_params *p, std::string &msg) {
    char buf[512];
    if (p->abi_version != VPBS_ABI_VERSION) {
        snprintf(buf, sizeof buf, "ABI version mismatch: got %d, library is %d", p->abi_version, VPBS_ABI_VERSION);
        msg = buf;
        return VPBS_ERR_INVALID;
    }
    if (p->precision != VPBS_F64 && p->precision != VPBS_F32) {
        msg = "precision must be VPBS_F64 or VPBS_F32";
        return VPBS_ERR_INVALID;
    }
    for (int d = 0; d < 3; d++) {
        if (!(p->Ls[d] > 0) || !std::isfinite(p->Ls[d])) {
            msg = "periods Ls must be positive and finite (open domains are not supported)";
            return VPBS_ERR_INVALID;
        }
        if (p->Ns[d] < 2) {
            msg = "Fourier grid Ns must be at least 2 in every direction";
            return VPBS_ERR_INVALID;
        }
    }
    if (!(p->alpha > 0) || !(p->rcut > 0)) {
        msg = "alpha and rcut must be positive";
        return VPBS_ERR_INVALID;
    }
    if (p->nquad < 1 || p->nquad > 16) {
        msg = "nquad must be in 1..16";
        return VPBS_ERR_INVALID;
    }
    if (p->nsubdiv < 1 || p->nsubdiv > 4) {
        msg = "nsubdiv must be in 1..4";
        return VPBS_ERR_INVALID;
    }
    if (p->nufft_m < 2 || p->nufft_m > 8) {
        msg = "nufft_m must be in 2..8";
        return VPBS_ERR_INVALID;
    }
    if (!(p->nufft_sigma > 1.0)) {
        msg = "nufft_sigma must be > 1";
        return VPBS_ERR_INVALID;
    }
    if (p->fft != VPBS_FFT_FUSED && p->fft != VPBS_FFT_ROCFFT && p->fft != VPBS_FFT_SEPARATE) {
        msg = "fft must be VPBS_FFT_FUSED (0), VPBS_FFT_ROCFFT (1) or VPBS_FFT_SEPARATE (2)";
        return VPBS_ERR_INVALID;
    }
    const double Lmin = std::fmin(p->Ls[0], std::fmin(p->Ls[1], p->Ls[2]));
    const int M = p->nsubdiv;
    const double rmax = ((double)M / (2 * M + 1)) * Lmin;  // CellLists.jl:115
    if (p->rcut > rmax) {
        snprintf(buf, sizeof buf,
                 "cutoff distance `rcut = %g` is larger than that allowed by the cell-list backend "
                 "(rcut ≤ M/(2M+1)·L = %g with M = %d)",
                 p->rcut, rmax, M);
        msg = buf;
        return VPBS_ERR_INVALID;
    }
    if (!(2 * p->rcut <= Lmin)) {  // params.jl:105-106
        msg = "cutoff distance rcut must be less than half the period in each direction";
        return VPBS_ERR_INVALID;
    }
    return VPBS_OK;
}

static void release(vpbs_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->s_sr) (void)hipStreamSynchronize(c->s_sr);
    if (c->s_lr) (void)hipStreamSynchronize(c->s_lr);
    longrange_release(c);
    learned_release(c);
    void *bufs[] = {c->nodes,     c->sp,        c->spp,      c->dlm,       c->dlp,        c->idx_prev,
                    c->points,    c->charges,   c->v_sr,     c->psi_sr,    c->v_lr,       c->psi_lr,
                    c->v_out,     c->psi_out,   c->src_key,  c->src_start, c->src_perm,   c->tgt_key,   c->tgt_cell,
                    c->tgt_start, c->tgt_perm,  c->src_pos,  c->src_chg,   c->src_idx,    c->stat_counts, c->stat_cands,
                    c->stat_hashes, c->lr_key,  c->lr_start, c->lr_active, c->lr_perm,   c->lr_i0,      c->lr_w,
                    c->sort_sr.rank, c->sort_sr.scan_tmp, c->sort_tg.rank, c->sort_tg.scan_tmp,
                    c->sort_lr.rank, c->sort_lr.scan_tmp, c->sort_lr.keys_out, c->sort_lr.radix_tmp,
                    c->sort_nd.rank, c->sort_nd.scan_tmp, c->sort_nd.keys_out, c->sort_nd.radix_tmp,
                    c->nd_key,    c->nd_start,  c->nd_perm,  c->nd_i0,    c->nd_rec,     c->slab_part,
                    c->fil_x, c->fil_ts, c->fil_c, c->fil_d, c->fil_e,
                    c->fil_fid, c->fil_n, c->fil_off, c->fil_T, c->fil_xoff, c->step_buf, c->solve_buf};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (int st = 0; st < ST_COUNT; st++) {
        if (c->ev_beg[st]) (void)hipEventDestroy(c->ev_beg[st]);
        if (c->ev_end[st]) (void)hipEventDestroy(c->ev_end[st]);
    }
    if (c->ev_lr_done) (void)hipEventDestroy(c->ev_lr_done);
    if (c->ev_nodes) (void)hipEventDestroy(c->ev_nodes);
    if (c->s_sr) (void)hipStreamDestroy(c->s_sr);
    if (c->s_lr) (void)hipStreamDestroy(c->s_lr);
    delete c;
}

template <typename T>
static int alloc(vpbs_ctx *c, T **p, int64_t n) {
    if (*p) {
        VPBS_HIP(c, hipFree(*p));
        *p = nullptr;
    }
    VPBS_HIP(c, hipMalloc((void **)p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    return VPBS_OK;
}

int ensure_capacity(vpbs_ctx *c, int64_t Np, int64_t Nq) {
    const size_t cs = c->p.precision == VPBS_F32 ? sizeof(float) : sizeof(double);
    if (Np > c->cap_np) {
        VPBS_TRY(alloc(c, &c->nodes, 3 * Np));
        VPBS_TRY(alloc(c, &c->sp, 3 * Np));
        VPBS_TRY(alloc(c, &c->spp, 3 * Np));
        VPBS_TRY(alloc(c, &c->dlm, Np));
        VPBS_TRY(alloc(c, &c->dlp, Np));
        VPBS_TRY(alloc(c, &c->idx_prev, Np));
        VPBS_TRY(alloc(c, &c->v_sr, 3 * Np));
        VPBS_TRY(alloc(c, &c->psi_sr, 3 * Np));
        VPBS_TRY(alloc(c, &c->v_lr, 3 * Np));
        VPBS_TRY(alloc(c, &c->psi_lr, 3 * Np));
        VPBS_TRY(alloc(c, &c->v_out, 3 * Np));
        VPBS_TRY(alloc(c, &c->psi_out, 3 * Np));
        VPBS_TRY(alloc(c, &c->tgt_key, Np));
        VPBS_TRY(alloc(c, &c->tgt_cell, 4 * Np));
        VPBS_TRY(alloc(c, &c->tgt_perm, Np));
        VPBS_TRY(alloc(c, &c->stat_counts, Np));
        VPBS_TRY(alloc(c, &c->stat_cands, Np));
        VPBS_TRY(alloc(c, &c->stat_hashes, Np));
        c->cap_np = Np;
    }
    if (Nq > c->cap_nq) {
        VPBS_TRY(alloc(c, &c->points, 3 * Nq));
        VPBS_TRY(alloc(c, &c->charges, 3 * Nq));
        VPBS_TRY(alloc(c, &c->src_key, Nq));
        VPBS_TRY(alloc(c, &c->src_perm, Nq));
        VPBS_TRY(alloc(c, &c->src_pos, 3 * Nq));
        VPBS_TRY(alloc(c, &c->src_idx, Nq));
        VPBS_TRY(alloc(c, &c->lr_key, Nq));
        VPBS_TRY(alloc(c, &c->lr_perm, Nq));
        VPBS_TRY(alloc(c, &c->lr_i0, 3 * Nq));
        if (c->src_chg) VPBS_HIP(c, hipFree(c->src_chg));
        if (c->lr_w) VPBS_HIP(c, hipFree(c->lr_w));
        c->src_chg = c->lr_w = nullptr;
        VPBS_HIP(c, hipMalloc(&c->src_chg, cs * 3 * (size_t)Nq));
        const size_t rec = (3 * 2 * (size_t)c->p.nufft_m + 3 + 3) & ~(size_t)3;  // rec_len(2m) in longrange.hip
        VPBS_HIP(c, hipMalloc(&c->lr_w, cs * rec * (size_t)Nq));
        c->cap_nq = Nq;
    }
    return VPBS_OK;
}

// Counts entries that would make the kernels index out of bounds: coordinates
// that are not finite or not folded into [0, L), and idx_prev outside [0, Np).
__global__ void count_invalid(int64_t n, const double *__restrict__ pos, Box box, const int32_t *__restrict__ idx,
                              int64_t Np, unsigned int *__restrict__ bad) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool ok = true;
    for (int d = 0; d < 3; d++) {
        double x = pos[d * n + i];
        ok = ok && (x >= 0.0) && (x < box.L[d]);  // false for NaN
    }
    if (idx) ok = ok && idx[i] >= 0 && idx[i] < Np;
    if (!ok) atomicAdd(bad, 1u);
}

static int host_check(vpbs_ctx *c, int64_t n, const double *pos, const int32_t *idx, int64_t Np, const char *what) {
    for (int64_t i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) {
            double x = pos[d * n + i];
            if (!(x >= 0.0 && x < c->p.Ls[d]))
                return fail(c, VPBS_ERR_INVALID,
                            "%s[%lld] component %d = %g is not folded into [0, L) (L = %g)", what, (long long)i, d,
                            x, c->p.Ls[d]);
        }
        if (idx && (idx[i] < 0 || idx[i] >= Np))
            return fail(c, VPBS_ERR_INVALID, "idx_prev[%lld] = %d out of range [0, %lld)", (long long)i, idx[i],
                        (long long)Np);
    }
    return VPBS_OK;
}

int device_check(vpbs_ctx *c) {
    unsigned int *bad = nullptr;
    VPBS_HIP(c, hipMalloc((void **)&bad, 2 * sizeof(unsigned int)));
    VPBS_HIP(c, hipMemset(bad, 0, 2 * sizeof(unsigned int)));
    count_invalid<<<(unsigned)((c->Np + 255) / 256), 256>>>(c->Np, c->nodes, c->box, c->idx_prev, c->Np, bad);
    count_invalid<<<(unsigned)((c->Nq + 255) / 256), 256>>>(c->Nq, c->points, c->box, nullptr, c->Np, bad + 1);
    unsigned int h[2] = {0, 0};
    VPBS_HIP(c, hipMemcpy(h, bad, sizeof h, hipMemcpyDeviceToHost));
    VPBS_HIP(c, hipFree(bad));
    if (h[0] || h[1])
        return fail(c, VPBS_ERR_INVALID,
                    "%u nodes and %u points are not folded into [0, L) or have idx_prev out of range", h[0], h[1]);
    return VPBS_OK;
}

static int set_points_impl(vpbs_ctx *c, int64_t Np, const double *nodes, const int32_t *idx_prev,
                           const double *sp, const double *spp, const double *dlm, const double *dlp, int64_t Nq,
                           const double *points, const double *charges, hipMemcpyKind kind) {
    if (Np < 1) return fail(c, VPBS_ERR_INVALID, "need at least one node (got Np = %lld)", (long long)Np);
    if (Nq != (int64_t)c->p.nquad * Np)
        return fail(c, VPBS_ERR_INVALID, "Nq = %lld must equal nquad * Np = %lld", (long long)Nq,
                    (long long)c->p.nquad * Np);
    if (Nq > INT32_MAX / 2) return fail(c, VPBS_ERR_INVALID, "too many quadrature points for 32-bit indices");
    if (!nodes || !idx_prev || !sp || !spp || !dlm || !dlp || !points || !charges)
        return fail(c, VPBS_ERR_INVALID, "null input array");
    VPBS_HIP(c, hipSetDevice(c->device));
    c->have_points = false;
    if (kind == hipMemcpyHostToDevice) {
        VPBS_TRY(host_check(c, Np, nodes, idx_prev, Np, "nodes"));
        VPBS_TRY(host_check(c, Nq, points, nullptr, Np, "points"));
    }
    VPBS_TRY(ensure_capacity(c, Np, Nq));
    const size_t d3 = sizeof(double) * 3 * (size_t)Np, d1 = sizeof(double) * (size_t)Np;
    VPBS_HIP(c, hipMemcpy(c->nodes, nodes, d3, kind));
    VPBS_HIP(c, hipMemcpy(c->idx_prev, idx_prev, sizeof(int32_t) * (size_t)Np, kind));
    VPBS_HIP(c, hipMemcpy(c->sp, sp, d3, kind));
    VPBS_HIP(c, hipMemcpy(c->spp, spp, d3, kind));
    VPBS_HIP(c, hipMemcpy(c->dlm, dlm, d1, kind));
    VPBS_HIP(c, hipMemcpy(c->dlp, dlp, d1, kind));
    VPBS_HIP(c, hipMemcpy(c->points, points, sizeof(double) * 3 * (size_t)Nq, kind));
    VPBS_HIP(c, hipMemcpy(c->charges, charges, sizeof(double) * 3 * (size_t)Nq, kind));
    c->Np = Np;
    c->Nq = Nq;
    c->timings.point_data = 0.0;
    c->local_ns = 0;  // no filament description on the device: local integrals stay with the caller
    c->have_filaments = false;
    if (kind == hipMemcpyDeviceToDevice) VPBS_TRY(device_check(c));
    c->have_points = true;
    c->points_version++;
    c->have_stats = false;
    c->last_vel = c->last_psi = false;
    return VPBS_OK;
}

}  // namespace vpbs

using namespace vpbs;

extern "C" {

int vpbs_create(vpbs_ctx **out, const vpbs_params *p) {
    if (!out || !p) return fail(nullptr, VPBS_ERR_INVALID, "null argument to vpbs_create");
    *out = nullptr;
    std::string msg;
    int rc = validate(p, msg);
    if (rc != VPBS_OK) return fail(nullptr, rc, "%s", msg.c_str());
    vpbs_ctx *c = new (std::nothrow) vpbs_ctx();
    if (!c) return fail(nullptr, VPBS_ERR_HIP, "out of host memory");
    c->p = *p;
    c->device = p->device;
    // VPBS_PAIR=tile selects the tiled pair kernel (one wavefront per tile of cells; for
    // comparisons: 6.3 ms against 1.3 ms for one wavefront per target at 1M nodes, FP32)
    if (const char *e = std::getenv("VPBS_PAIR")) c->pair_tiled = std::strcmp(e, "tile") == 0;
    for (int d = 0; d < 3; d++) {
        c->box.L[d] = p->Ls[d];
        c->box.Lh[d] = p->Ls[d] / 2;
        c->Nc[d] = (int32_t)std::floor((p->Ls[d] * (double)p->nsubdiv) / p->rcut);  // CellLists.jl:132-137
    }
    c->ncells = (int64_t)c->Nc[0] * c->Nc[1] * c->Nc[2];
    auto bail = [&](int code) {
        g_create_error = c->err;
        release(c);
        return code;
    };
    if (c->ncells > INT32_MAX / 2) {
        fail(c, VPBS_ERR_INVALID, "too many cells (%lld)", (long long)c->ncells);
        return bail(VPBS_ERR_INVALID);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        fail(c, VPBS_ERR_HIP, "no HIP device available");
        return bail(VPBS_ERR_HIP);
    }
    if (p->device < 0 || p->device >= ndev) {
        fail(c, VPBS_ERR_INVALID, "device %d out of range (%d devices)", p->device, ndev);
        return bail(VPBS_ERR_INVALID);
    }
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->s_sr, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->s_lr, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_lr_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_nodes, hipEventDisableTiming) != hipSuccess) {
        fail(c, VPBS_ERR_HIP, "stream/event creation failed");
        return bail(VPBS_ERR_HIP);
    }
    for (int st = 0; st < ST_COUNT; st++) {
        if (hipEventCreate(&c->ev_beg[st]) != hipSuccess || hipEventCreate(&c->ev_end[st]) != hipSuccess) {
            fail(c, VPBS_ERR_HIP, "event creation failed");
            return bail(VPBS_ERR_HIP);
        }
    }
    if (hipMalloc((void **)&c->src_start, sizeof(int32_t) * (size_t)(c->ncells + 1)) != hipSuccess ||
        hipMalloc((void **)&c->tgt_start, sizeof(int32_t) * (size_t)(c->ncells + 1)) != hipSuccess) {
        fail(c, VPBS_ERR_HIP, "cell-list allocation failed");
        return bail(VPBS_ERR_HIP);
    }
    if (rocfft_setup() != rocfft_status_success) {
        fail(c, VPBS_ERR_FFT, "rocfft_setup failed");
        return bail(VPBS_ERR_FFT);
    }
    rc = longrange_setup(c);
    if (rc != VPBS_OK) return bail(rc);
    if (hipMalloc((void **)&c->lr_start, sizeof(int32_t) * (size_t)(c->ntiles + 1)) != hipSuccess ||
        (c->slab_interp && hipMalloc((void **)&c->nd_start, sizeof(int32_t) * (size_t)(c->nslabs + 1)) != hipSuccess) ||
        hipMalloc((void **)&c->lr_active, (size_t)c->ntiles) != hipSuccess) {
        fail(c, VPBS_ERR_HIP, "tile allocation failed");
        return bail(VPBS_ERR_HIP);
    }
    *out = c;
    return VPBS_OK;
}

void vpbs_destroy(vpbs_ctx *c) { release(c); }

const char *vpbs_last_error(const vpbs_ctx *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int vpbs_set_points(vpbs_ctx *c, int64_t Np, const double *nodes, const int32_t *idx_prev, const double *sp,
                    const double *spp, const double *dlm, const double *dlp, int64_t Nq, const double *points,
                    const double *charges) {
    if (!c) return VPBS_ERR_INVALID;
    return set_points_impl(c, Np, nodes, idx_prev, sp, spp, dlm, dlp, Nq, points, charges, hipMemcpyHostToDevice);
}

int vpbs_set_points_device(vpbs_ctx *c, int64_t Np, const double *nodes, const int32_t *idx_prev,
                           const double *sp, const double *spp, const double *dlm, const double *dlp, int64_t Nq,
                           const double *points, const double *charges) {
    if (!c) return VPBS_ERR_INVALID;
    return set_points_impl(c, Np, nodes, idx_prev, sp, spp, dlm, dlp, Nq, points, charges,
                           hipMemcpyDeviceToDevice);
}

int vpbs_set_filaments(vpbs_ctx *c, int64_t Nf, const int64_t *counts, const double *X, const double *knots,
                       const double *periods, const double *offsets, const double *coefs, const double *quad_nodes,
                       const double *quad_weights, double lia_segment_fraction, int32_t nquad_ns,
                       const double *quad_ns_nodes, const double *quad_ns_weights) {
    if (!c) return VPBS_ERR_INVALID;
    return build_point_data(c, Nf, counts, X, knots, periods, offsets, coefs, quad_nodes, quad_weights,
                            lia_segment_fraction, nquad_ns, quad_ns_nodes, quad_ns_weights);
}

int vpbs_get_points(vpbs_ctx *c, double *nodes, int32_t *idx_prev, double *sp, double *spp, double *dlm,
                    double *dlp, double *points, double *charges) {
    if (!c) return VPBS_ERR_INVALID;
    if (!c->have_points) return fail(c, VPBS_ERR_STATE, "no point data: call vpbs_set_points or vpbs_set_filaments");
    VPBS_HIP(c, hipSetDevice(c->device));
    const size_t d3 = sizeof(double) * 3 * (size_t)c->Np, d1 = sizeof(double) * (size_t)c->Np;
    const size_t q3 = sizeof(double) * 3 * (size_t)c->Nq;
    const auto D2H = hipMemcpyDeviceToHost;
    if (nodes) VPBS_HIP(c, hipMemcpy(nodes, c->nodes, d3, D2H));
    if (idx_prev) VPBS_HIP(c, hipMemcpy(idx_prev, c->idx_prev, sizeof(int32_t) * (size_t)c->Np, D2H));
    if (sp) VPBS_HIP(c, hipMemcpy(sp, c->sp, d3, D2H));
    if (spp) VPBS_HIP(c, hipMemcpy(spp, c->spp, d3, D2H));
    if (dlm) VPBS_HIP(c, hipMemcpy(dlm, c->dlm, d1, D2H));
    if (dlp) VPBS_HIP(c, hipMemcpy(dlp, c->dlp, d1, D2H));
    if (points) VPBS_HIP(c, hipMemcpy(points, c->points, q3, D2H));
    if (charges) VPBS_HIP(c, hipMemcpy(charges, c->charges, q3, D2H));
    return VPBS_OK;
}

int vpbs_compute(vpbs_ctx *c, uint32_t flags, double integration_cutoff, double *v_out, double *psi_out) {
    if (!c) return VPBS_ERR_INVALID;
    if (!c->have_points) return fail(c, VPBS_ERR_STATE, "vpbs_compute called before vpbs_set_points");
    const bool vel = flags & VPBS_VELOCITY, psi = flags & VPBS_STREAMFUNCTION;
    if (!vel && !psi) return fail(c, VPBS_ERR_INVALID, "request VPBS_VELOCITY and/or VPBS_STREAMFUNCTION");
    if ((flags & VPBS_LIA_OFF) && (flags & VPBS_LIA_ONLY))
        return fail(c, VPBS_ERR_INVALID, "VPBS_LIA_OFF and VPBS_LIA_ONLY are exclusive");
    if (!std::isnan(integration_cutoff) && !(integration_cutoff > 0))
        return fail(c, VPBS_ERR_INVALID, "integration_cutoff must be positive (or NaN for the default)");
    VPBS_HIP(c, hipSetDevice(c->device));
    // Both streams are drained on every return, the early (error) returns included, so that
    // no kernel of this call still reads the point data when the caller replaces it.
    StreamDrain drain{c};
    auto t0 = std::chrono::steady_clock::now();
    for (int st = 0; st < ST_POINTS; st++) c->stage_ran[st] = false;
    const bool with_sr = !(flags & VPBS_NO_SHORTRANGE);
    const bool with_lr = !(flags & VPBS_NO_LONGRANGE) && !(flags & VPBS_LIA_ONLY);
    const int64_t n3 = 3 * c->Np;

    // learned evaluation (vpbs_compute_learned): both ranges are present
    const bool learned = c->lc_active && with_lr && with_sr && vel;
    if (with_lr)
        VPBS_TRY(longrange_compute(c, flags));
    else
        c->uhat_quantity = 0;
    if (learned) VPBS_TRY(learned_check_long(c));
    if (with_sr) {
        VPBS_TRY(shortrange_compute(c, flags, integration_cutoff));
        // BiotSavart.jl:570-576: with the short range, unless LIA = Val(:only)
        if (!(flags & VPBS_LIA_ONLY)) VPBS_TRY(add_local_integrals(c, flags, c->s_sr));
        // features and inference beside the long range
        if (learned) VPBS_TRY(learned_after_shortrange(c));
    }
    if (learned) VPBS_TRY(learned_check_short(c));
    if (with_lr) VPBS_HIP(c, hipEventRecord(c->ev_lr_done, c->s_lr));

    // merge on the short-range stream
    hipStream_t s = c->s_sr;
    if (with_lr) VPBS_HIP(c, hipStreamWaitEvent(s, c->ev_lr_done, 0));
    stage_begin(c, ST_MERGE, s);
    struct F {
        bool on;
        double *sr, *lr, *out;
    } fs[2] = {{vel, c->v_sr, c->v_lr, c->v_out}, {psi, c->psi_sr, c->psi_lr, c->psi_out}};
    for (auto &f : fs) {
        if (!f.on) continue;
        if (with_sr && with_lr)
            VPBS_TRY(merge_outputs(c, f.sr, f.lr, f.out, n3, s));
        else if (with_sr)
            VPBS_HIP(c, hipMemcpyAsync(f.out, f.sr, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
        else if (with_lr)
            VPBS_HIP(c, hipMemcpyAsync(f.out, f.lr, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
        else
            VPBS_HIP(c, hipMemsetAsync(f.out, 0, sizeof(double) * n3, s));
    }
    stage_end(c, ST_MERGE, s);
    if (learned) VPBS_TRY(learned_after_merge(c));
    VPBS_HIP(c, hipStreamSynchronize(s));
    VPBS_HIP(c, hipStreamSynchronize(c->s_lr));
    VPBS_HIP(c, hipGetLastError());
    c->last_vel = vel;
    c->last_psi = psi;
    auto t1 = std::chrono::steady_clock::now();

    collect_timings(c);
    c->timings.total = std::chrono::duration<double, std::milli>(t1 - t0).count();
    vpbs_timings &T = c->timings;
    T.pair_candidates = 0;
    if (with_sr && (flags & VPBS_NEIGHBOUR_STATS) && !(flags & VPBS_LIA_ONLY) && c->Np > 0) {
        std::vector<int32_t> nc((size_t)c->Np);
        VPBS_HIP(c, hipMemcpy(nc.data(), c->stat_cands, sizeof(int32_t) * (size_t)c->Np, hipMemcpyDeviceToHost));
        for (int32_t v : nc) T.pair_candidates += v;
    }
    for (int d = 0; d < 3; d++) {
        T.grid[d] = c->G[d];
        T.cells[d] = c->Nc[d];
    }
    if (v_out && vel) VPBS_HIP(c, hipMemcpy(v_out, c->v_out, sizeof(double) * n3, hipMemcpyDeviceToHost));
    if (psi_out && psi) VPBS_HIP(c, hipMemcpy(psi_out, c->psi_out, sizeof(double) * n3, hipMemcpyDeviceToHost));
    return VPBS_OK;
}

int vpbs_get_outputs(vpbs_ctx *c, double *v_out, double *psi_out) {
    if (!c) return VPBS_ERR_INVALID;
    VPBS_HIP(c, hipSetDevice(c->device));
    const size_t n3 = sizeof(double) * 3 * (size_t)c->Np;
    if (v_out) {
        if (!c->last_vel) return fail(c, VPBS_ERR_STATE, "velocity was not computed by the last call");
        VPBS_HIP(c, hipMemcpy(v_out, c->v_out, n3, hipMemcpyDeviceToHost));
    }
    if (psi_out) {
        if (!c->last_psi) return fail(c, VPBS_ERR_STATE, "streamfunction was not computed by the last call");
        VPBS_HIP(c, hipMemcpy(psi_out, c->psi_out, n3, hipMemcpyDeviceToHost));
    }
    return VPBS_OK;
}

int vpbs_output_device_ptrs(vpbs_ctx *c, const double **v_dev, const double **psi_dev) {
    if (!c) return VPBS_ERR_INVALID;
    if (v_dev) *v_dev = c->last_vel ? c->v_out : nullptr;
    if (psi_dev) *psi_dev = c->last_psi ? c->psi_out : nullptr;
    return VPBS_OK;
}

int vpbs_neighbour_stats(vpbs_ctx *c, int64_t *counts, uint64_t *hashes) {
    if (!c) return VPBS_ERR_INVALID;
    if (!c->have_stats)
        return fail(c, VPBS_ERR_STATE, "no neighbour statistics: run vpbs_compute with VPBS_NEIGHBOUR_STATS");
    VPBS_HIP(c, hipSetDevice(c->device));
    if (counts) VPBS_HIP(c, hipMemcpy(counts, c->stat_counts, sizeof(int64_t) * c->Np, hipMemcpyDeviceToHost));
    if (hashes) VPBS_HIP(c, hipMemcpy(hashes, c->stat_hashes, sizeof(uint64_t) * c->Np, hipMemcpyDeviceToHost));
    return VPBS_OK;
}

int vpbs_advance(vpbs_ctx *c, double dt, int32_t nsteps) {
    if (!c) return VPBS_ERR_INVALID;
    if (!c->have_filaments) return fail(c, VPBS_ERR_STATE, "no filaments on the device: call vpbs_set_filaments first");
    if (!(dt > 0) || !std::isfinite(dt)) return fail(c, VPBS_ERR_INVALID, "time step must be positive and finite");
    if (nsteps < 0) return fail(c, VPBS_ERR_INVALID, "number of steps must be >= 0");
    VPBS_HIP(c, hipSetDevice(c->device));
    return advance_filaments(c, dt, nsteps);
}

int vpbs_get_filaments(vpbs_ctx *c, double *X, double *knots, double *periods, double *coefs) {
    if (!c) return VPBS_ERR_INVALID;
    if (!c->have_filaments) return fail(c, VPBS_ERR_STATE, "no filaments on the device: call vpbs_set_filaments first");
    VPBS_HIP(c, hipSetDevice(c->device));
    const size_t d3 = sizeof(double) * 3 * (size_t)c->Np;
    if (X) VPBS_HIP(c, hipMemcpy(X, c->fil_x, d3, hipMemcpyDeviceToHost));
    if (coefs) VPBS_HIP(c, hipMemcpy(coefs, c->fil_c, d3, hipMemcpyDeviceToHost));
    if (knots) VPBS_HIP(c, hipMemcpy(knots, c->fil_ts, sizeof(double) * (size_t)c->Np, hipMemcpyDeviceToHost));
    if (periods) VPBS_HIP(c, hipMemcpy(periods, c->fil_T, sizeof(double) * (size_t)c->pd_nf, hipMemcpyDeviceToHost));
    return VPBS_OK;
}

int vpbs_get_uhat(vpbs_ctx *c, int32_t *quantity, int64_t *dims3, double *out) {
    if (!c) return VPBS_ERR_INVALID;
    VPBS_HIP(c, hipSetDevice(c->device));
    return longrange_get_uhat(c, quantity, dims3, out);
}

int vpbs_energy_spectrum(vpbs_ctx *c, int64_t nk, double dk, double *Ek) {
    if (!c) return VPBS_ERR_INVALID;
    VPBS_HIP(c, hipSetDevice(c->device));
    return longrange_spectrum(c, nk, dk, Ek);
}

int vpbs_energy_spectrum_noalloc(vpbs_ctx *c, int64_t nk, double dk, double *Ek) {
    if (!c) return VPBS_ERR_INVALID;
    VPBS_HIP(c, hipSetDevice(c->device));
    return longrange_spectrum_noalloc(c, nk, dk, Ek);
}

int vpbs_interpolate_band(vpbs_ctx *c, double kmin, double kmax, double *v_out) {
    if (!c) return VPBS_ERR_INVALID;
    if (!c->have_points) return fail(c, VPBS_ERR_STATE, "no point data: call vpbs_set_points or vpbs_set_filaments");
    VPBS_HIP(c, hipSetDevice(c->device));
    return longrange_band(c, kmin, kmax, v_out);
}

int vpbs_get_timings(vpbs_ctx *c, vpbs_timings *t) {
    if (!c || !t) return VPBS_ERR_INVALID;
    *t = c->timings;
    return VPBS_OK;
}

int vpbs_get_sizes(vpbs_ctx *c, int64_t *grid3, int64_t *cells3) {
    if (!c) return VPBS_ERR_INVALID;
    for (int d = 0; d < 3; d++) {
        if (grid3) grid3[d] = c->G[d];
        if (cells3) cells3[d] = c->Nc[d];
    }
    return VPBS_OK;
}

}  // extern "C"
