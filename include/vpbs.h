/*
 * vpbs.h — C ABI of libvpbs, the MI355X (gfx950) Ewald-split Biot–Savart
 * library. It replaces, behind VortexPasta's backend plug-in interface, the
 * short-range backend (CellListsBackend) and the long-range backend
 * (NonuniformFFTsBackend) of BiotSavart.compute_on_nodes!.
 *
 * All array arguments are plain pointers with explicit sizes; vectors are
 * structure-of-arrays, component-major: a field of n 3-vectors is 3n doubles
 * laid out as x[0..n), y[0..n), z[0..n) — the memory layout of the
 * reference's StructVector{Vec3{T}} (src/BiotSavart/pointdata.jl:21-39).
 * Inputs and outputs are always FP64; the internal arithmetic is chosen by
 * vpbs_params.precision.
 *
 * Ownership: the caller owns every host array and only lends it for the
 * duration of a call. The library owns all device buffers, FFT plans and
 * streams, and reuses them across calls. Every call returns VPBS_OK (0) or a
 * negative error code; vpbs_last_error() gives the message. A context is
 * bound to one GPU and calls on one context must not overlap.
 */
#ifndef VPBS_H
#define VPBS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VPBS_ABI_VERSION 1

/* error codes */
#define VPBS_OK 0
#define VPBS_ERR_INVALID -1  /* invalid argument or parameter */
#define VPBS_ERR_HIP -2      /* HIP runtime error (allocation, launch, copy) */
#define VPBS_ERR_FFT -3      /* rocFFT error */
#define VPBS_ERR_STATE -4    /* call out of order (e.g. compute before set_points) */

/* precision of the bulk arithmetic */
#define VPBS_F64 0 /* plain FP64 path */
#define VPBS_F32 1 /* FP32 bulk, FP64 node differencing / LIA / local self terms */

/* transforms of the long-range part (vpbs_params.fft) */
#define VPBS_FFT_FUSED 0    /* hand-written FFTs fused with the pencil reorderings and the Fourier-space kernel;
                               on x-grids of 192, 384 or 768 cells (768: FP32 only) also the inverse x
                               transform fused with the interpolation (no real grid written on that side) */
#define VPBS_FFT_ROCFFT 1   /* rocFFT 1D transforms with separate reordering and Fourier-space kernels */
#define VPBS_FFT_SEPARATE 2 /* the hand-written FFTs of VPBS_FFT_FUSED with the interpolation in a kernel of its own */

/* vpbs_compute flags */
#define VPBS_VELOCITY 1u
#define VPBS_STREAMFUNCTION 2u
#define VPBS_LIA_OFF 4u        /* LIA = Val(false) */
#define VPBS_LIA_ONLY 8u       /* LIA = Val(:only) */
#define VPBS_NO_SHORTRANGE 16u /* shortrange = false */
#define VPBS_NO_LONGRANGE 32u  /* longrange = false */
#define VPBS_NEIGHBOUR_STATS 64u /* record per-node accepted-pair counts and hashes */
#define VPBS_KEEP_UHAT 128u      /* keep the unsmoothed long-range field in Fourier space (see vpbs_get_uhat) */

/* quantity held by the kept Fourier field (LongRangeCacheState.quantity) */
#define VPBS_UHAT_NONE 0
#define VPBS_UHAT_VELOCITY 1
#define VPBS_UHAT_STREAMFUNCTION 2

/*
 * Parameters. Mirrors ParamsBiotSavart (src/BiotSavart/params.jl:326-376)
 * with the Gaussian splitting already resolved (splitting/gaussian.jl:109-200):
 * alpha, rcut and Ns are the derived values.
 */
typedef struct vpbs_params {
    int32_t abi_version;   /* must be VPBS_ABI_VERSION */
    int32_t precision;     /* VPBS_F64 or VPBS_F32 */
    double Gamma;          /* circulation Γ */
    double a;              /* core size */
    double Delta;          /* core parameter Δ */
    double Ls[3];          /* periods */
    double alpha;          /* Ewald splitting parameter */
    double rcut;           /* short-range cut-off distance */
    int32_t Ns[3];         /* Fourier grid (retained modes follow longrange.jl:556-583) */
    int32_t nquad;         /* quadrature points per segment */
    int32_t nsubdiv;       /* cell-list subdivisions M (CellListsBackend(M)) */
    int32_t nufft_m;       /* NUFFT kernel half-support (HalfSupport(m)) */
    double nufft_sigma;    /* NUFFT oversampling factor σ */
    int32_t avoid_explicit_erf; /* params.jl:289-298 */
    int32_t truncate_spherical; /* params.jl:243-246 */
    int32_t device;        /* HIP device ordinal */
    int32_t fft;           /* VPBS_FFT_FUSED (default), VPBS_FFT_ROCFFT or VPBS_FFT_SEPARATE; grids the
                              hand-written transforms do not cover (x, y > 6144 FP32 / 3072 FP64,
                              z > 3072 / 1536) use rocFFT */
} vpbs_params;

/* Per-stage GPU times of the last vpbs_compute, in milliseconds (HIP events). */
typedef struct vpbs_timings {
    double total;        /* whole compute, host wall-clock (includes waits) */
    double local_term;   /* LIA kernel */
    double cell_build;   /* source + target counting sorts */
    double pair;         /* short-range pair kernel */
    double self;         /* self-interaction removal */
    double lr_sort;      /* NUFFT tile sort of quadrature points */
    double spread;       /* type-1 spreading (all 3 components) */
    double fft_forward;  /* x and y transforms (fused: with the pencil reorderings), 3 components */
    double fourier;      /* Fourier-space kernel (fused: with the forward and inverse z transforms) */
    double fft_backward; /* inverse y and x transforms of the velocity (only y when the interpolation
                            runs the x transform) */
    double interp;       /* type-2 interpolation (with the inverse x transform when fused; with the
                            streamfunction, also its inverse transforms) */
    double merge;        /* SR + LR sum */
    int64_t pair_candidates; /* candidate sources the pair kernel tested (sum over nodes of the sources in
                                the cells of the (2M+1)^3 neighbourhood that the rcut sphere around the
                                node reaches); set by calls with VPBS_NEIGHBOUR_STATS, else 0 */
    int64_t grid[3];     /* oversampled NUFFT grid */
    int64_t cells[3];    /* cell-list cells per direction */
    double point_data;   /* kernels of the last vpbs_set_filaments (0 after vpbs_set_points*) */
    double node_sort;    /* slab sort and kernel-weight records of the nodes for the interpolation (runs on
                            the short-range stream ahead of the pair sum, beside spreading) */
    double features;     /* learned correction: node cell list and per-node features */
    double mlp;          /* learned correction: network inference and the corrected sum */
    double check;        /* learned correction: sampled check (converged-setting difference at the sample) */
} vpbs_timings;

typedef struct vpbs_ctx vpbs_ctx;

/*
 * Create a context: validates parameters (rcut against the cell-list limit
 * M/(2M+1)·L, CellLists.jl:115,143-147; 2·rcut ≤ min(L), params.jl:105-106),
 * builds FFT plans and deconvolution tables.
 * Replaces init_cache_short (shortrange/backends/cell_lists.jl:103-115) and
 * init_cache_long_ewald (longrange/backends/nonuniformffts.jl:153-171).
 */
int vpbs_create(vpbs_ctx **ctx, const vpbs_params *params);

/* Release every resource of the context. */
void vpbs_destroy(vpbs_ctx *ctx);

/* Message of the last failed call on ctx (or of the last failed vpbs_create when ctx is NULL). */
const char *vpbs_last_error(const vpbs_ctx *ctx);

/*
 * Upload point data (host pointers). Replaces the H2D copies of
 * do_shortrange!/do_longrange! (src/BiotSavart/BiotSavart.jl:425-430,483-499)
 * of the PointData built by add_point_charges! (pointdata_impl.jl:121-188):
 *   nodes[3*Np]    folded filament nodes
 *   idx_prev[Np]   0-based index of the previous node on the same filament
 *   sp, spp[3*Np]  s′ and s″ on nodes
 *   dlm, dlp[Np]   sub-segment lengths δ⁻, δ⁺
 *   points[3*Nq], charges[3*Nq] quadrature points (folded) and w·Δt·s′;
 *   Nq must equal nquad·Np, points of segment i at nquad·i .. nquad·i+nquad-1.
 */
int vpbs_set_points(vpbs_ctx *ctx, int64_t Np, const double *nodes, const int32_t *idx_prev, const double *sp,
                    const double *spp, const double *dlm, const double *dlp, int64_t Nq, const double *points,
                    const double *charges);

/* Same as vpbs_set_points with device pointers (device-resident pipelines; device-to-device copies). */
int vpbs_set_points_device(vpbs_ctx *ctx, int64_t Np, const double *nodes, const int32_t *idx_prev,
                           const double *sp, const double *spp, const double *dlm, const double *dlp, int64_t Nq,
                           const double *points, const double *charges);

/*
 * Build the point data on the device from the filaments' spline description:
 * the device counterpart of add_point_charges! (pointdata_impl.jl:121-219),
 * replacing the host loop and the H2D copy of its ~240 B/node output by an
 * upload of ~56 B/node. The result is what vpbs_set_points would upload, and
 * is checked the same way (folded into [0, L), idx_prev in range).
 *   Nf                number of filaments; counts[Nf] nodes of each (≥ 5), Np = Σ counts;
 *                     nodes of filament f are f's block of the linear node index
 *   X[3*Np]           node positions, not necessarily folded
 *   knots[Np]         spline parameter t_i of each node, t = 0 at each filament's
 *                     first node (chordal knots, spline.jl); periods[Nf] knot period T_f
 *   offsets[3*Nf]     end-to-end offset of each filament (0 when closed; periodic lines)
 *   coefs[3*Np]       B-spline coefficients of the periodised curve X(t) - (t/T)·offset
 *                     (QuinticSplineMethod, Filaments/spline/quintic.jl; index j ↔ B_{j-5,6})
 *   quad_nodes, quad_weights[nquad]  quadrature rule on [0, 1] (GaussLegendre);
 *                     both NULL for NoQuadrature (straight segments, nquad = 1)
 *   lia_segment_fraction  NaN for `nothing`, else γ ∈ (0, 1]: δ± then cover the
 *                     fraction γ of each segment (local_integrals.jl:100-101)
 *   nquad_ns, quad_ns_nodes, quad_ns_weights  quadrature_near_singularity rule on [0, 1]
 *                     (≤ 16 points). With lia_segment_fraction set and nquad_ns > 0,
 *                     vpbs_compute also adds the near-singular local integrals of
 *                     add_local_integrals! (local_integrals.jl:209-305) on the device,
 *                     in FP64, whenever the short range runs without VPBS_LIA_ONLY
 *                     (BiotSavart.jl:570-576); otherwise they stay with the caller.
 */
int vpbs_set_filaments(vpbs_ctx *ctx, int64_t Nf, const int64_t *counts, const double *X, const double *knots,
                       const double *periods, const double *offsets, const double *coefs, const double *quad_nodes,
                       const double *quad_weights, double lia_segment_fraction, int32_t nquad_ns,
                       const double *quad_ns_nodes, const double *quad_ns_weights);

/*
 * Advance the filaments uploaded by vpbs_set_filaments by nsteps steps of the
 * SSPRK33 scheme with their Biot–Savart velocity (Timestepping.step!,
 * Timestepping.jl:1038-1076; SSPRK33.jl:17-39; advect.jl:20-72), entirely on
 * the device: each stage builds the point data, evaluates the velocity
 * (VPBS_VELOCITY, LIA on, short and long range) and moves the nodes; within a
 * step the knots are kept and only the spline coefficients are solved again;
 * after the step the knots are recomputed from the new nodes (chordal). No
 * forcing, reconnections or adaptive time step. The point data of the final
 * state is left in the context. When a network is set up (vpbs_learned_setup
 * with a non-NULL net), every evaluation is the learned one (vpbs_compute_learned:
 * cheap splitting, correction, sampled check and fallback); see vpbs_learned_stats.
 */
int vpbs_advance(vpbs_ctx *ctx, double dt, int32_t nsteps);

/* Filament state on the device (after vpbs_set_filaments / vpbs_advance): nodes X[3*Np]
 * (not folded), knots[Np], periods[Nf], spline coefficients coefs[3*Np]; any may be NULL. */
int vpbs_get_filaments(vpbs_ctx *ctx, double *X, double *knots, double *periods, double *coefs);

/* Copy the device point data (from vpbs_set_points* or vpbs_set_filaments) to host arrays; any may be NULL. */
int vpbs_get_points(vpbs_ctx *ctx, double *nodes, int32_t *idx_prev, double *sp, double *spp, double *dlm,
                    double *dlp, double *points, double *charges);

/*
 * Compute velocity and/or streamfunction on nodes: the device part of
 * _compute_on_nodes! (BiotSavart.jl:515-612) — LIA term, short-range pair sum,
 * self-interaction removal and long-range NUFFT path, summed. The host-side
 * background-vorticity correction of ψ (BiotSavart.jl:255-286) stays with the
 * caller, and so do the lia_segment_fraction integrals unless the point data
 * came from vpbs_set_filaments with a near-singularity rule.
 * integration_cutoff: NaN for the default exp(Δ)a/2 / exp(Δ-1)a/2 cut-offs
 * (local_integrals.jl:115-125). v_out / psi_out: host arrays of 3*Np doubles,
 * or NULL to leave the result on the device (see vpbs_get_outputs).
 */
int vpbs_compute(vpbs_ctx *ctx, uint32_t flags, double integration_cutoff, double *v_out, double *psi_out);

/* Copy the outputs of the last vpbs_compute to host arrays (either may be NULL). */
int vpbs_get_outputs(vpbs_ctx *ctx, double *v_out, double *psi_out);

/* Device pointers of the last outputs (3*Np doubles each; NULL if not computed). Valid until the next call. */
int vpbs_output_device_ptrs(vpbs_ctx *ctx, const double **v_dev, const double **psi_dev);

/*
 * Per-node short-range neighbour statistics of the last compute run with
 * VPBS_NEIGHBOUR_STATS: number of accepted (node, quadrature point) pairs
 * (r² ≤ rcut²) and the wrapping sum of splitmix64(j) over accepted j.
 * Bit-exact comparison target for the cell-list neighbour sets.
 */
int vpbs_neighbour_stats(vpbs_ctx *ctx, int64_t *counts, uint64_t *hashes);

/*
 * Long-range field in Fourier space kept by the last vpbs_compute run with
 * VPBS_KEEP_UHAT: replaces get_longrange_field_fourier (cache.jl:77-83).
 * Like the reference's cache after compute_on_nodes!, it holds the unsmoothed
 * velocity v̂ = i k × ω̂/k² (or the streamfunction ψ̂ = ω̂/k² when only ψ was
 * computed; longrange.jl:102-224), without the Ewald factor. quantity:
 * VPBS_UHAT_*; dims3 = (N0/2+1, N1, N2); out (may be NULL): 3 components, each
 * dims3 complex values as (re, im) FP64 pairs with kx fastest (the reference's
 * uhat layout); modes outside the retained set are zero.
 */
int vpbs_get_uhat(vpbs_ctx *ctx, int32_t *quantity, int64_t *dims3, double *out);

/*
 * Energy spectrum of the kept field on the device: Diagnostics.energy_spectrum!
 * (Diagnostics/spectra.jl:89-196). Bins k_n = n·dk, n < nk ≤ 8192; mode k adds
 * f·|v̂|²/dk to bin trunc(|k|/dk + 1/2), f = 1/2 on the kx = 0 plane, else 1.
 */
int vpbs_energy_spectrum(vpbs_ctx *ctx, int64_t nk, double dk, double *Ek);

/*
 * vpbs_energy_spectrum without any allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * a time-stepping or serving loop that forbids hipMalloc). The per-block partial
 * spectra live in a fixed device array of 256 x 1024 + 1024 doubles (2.1 MB)
 * that each GPU holds until the process ends; nk must be at most 1024 (larger
 * nk: VPBS_ERR_INVALID). Ek: nk doubles, host or device memory. Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (several
 * contexts, several host threads) race and must be serialised by the caller.
 */
int vpbs_energy_spectrum_noalloc(vpbs_ctx *ctx, int64_t nk, double dk, double *Ek);

/*
 * Kept field restricted to kmin² ≤ k² ≤ kmax², interpolated at the current
 * nodes (type-2 NUFFT): the band velocity behind Forcing.FourierBandForcingBS
 * (fourier_band_bs.jl; band as SyntheticFields.set_wavevector_band!) and, with
 * kmin = 0, the low-pass velocity of SmallScaleDissipationBS
 * (dissipation_bs_small_scales.jl:109-128). v_out: 3*Np doubles (host).
 */
int vpbs_interpolate_band(vpbs_ctx *ctx, double kmin, double kmax, double *v_out);

/* Per-stage timings of the last vpbs_compute. */
int vpbs_get_timings(vpbs_ctx *ctx, vpbs_timings *t);

/* Oversampled grid and cell counts chosen by the context (for reporting). */
int vpbs_get_sizes(vpbs_ctx *ctx, int64_t *grid3, int64_t *cells3);

/*
 * ---------------------------------------------------------------------------
 * Learned correction. No counterpart in the reference: a small network
 * predicts, per node, the difference between the velocity of a cheap splitting
 * (small β: short cut-off, coarse grid) and that of a converged one (β = 4 on
 * the same grid), from local geometric features. The plain path stays the
 * reference: every learned evaluation checks the corrected velocity on a sample of
 * nodes against the converged setting and falls back to it when the error is
 * above tolerance.
 *
 * Network (vpbs_mlp): x (n_in) → [Linear → SiLU] × n_layers (width n_hidden) →
 * Linear → y (n_out). Parameters are one flat FP32 array: for each layer
 * l = 0..n_layers the weights W_l[out][in] (row-major) then the bias b_l[out].
 * Inputs are normalised as (x − in_shift)·in_scale and the output is
 * multiplied by out_scale. Training (FP32 forward, backward and Adam kernels)
 * minimises (1/B) Σ_b Σ_c (y_bc − t_bc)²/out_scale_c² over batches of B
 * samples. Inference runs on BF16 MFMA with FP32 accumulation; it is compiled
 * for (n_in, n_hidden, n_layers) = (128, 128, 3), (128, 128, 2) and (32, 32, 2).
 * Sample arrays X (n rows of n_in floats), T and Y (n rows of n_out floats)
 * are DEVICE pointers; parameter arrays are host pointers.
 */
typedef struct vpbs_mlp vpbs_mlp;

/* n_in and n_hidden multiples of 32, 1 ≤ n_layers ≤ 7, 1 ≤ n_out ≤ 16. */
int vpbs_mlp_create(vpbs_mlp **net, int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out,
                    int32_t device);
void vpbs_mlp_destroy(vpbs_mlp *net);
const char *vpbs_mlp_last_error(const vpbs_mlp *net);
int64_t vpbs_mlp_num_params(const vpbs_mlp *net);
/* Set parameters and normalisation (resets the Adam state); get them back (any may be NULL). */
int vpbs_mlp_set_params(vpbs_mlp *net, const float *params, const float *in_shift, const float *in_scale,
                        const float *out_scale);
int vpbs_mlp_get_params(vpbs_mlp *net, float *params, float *in_shift, float *in_scale, float *out_scale);
/* FP32 forward pass (the training arithmetic): Y = net(X), n ≤ 2^24. */
int vpbs_mlp_forward(vpbs_mlp *net, int64_t n, const float *X, float *Y);
/* BF16-MFMA inference: Y = net(X). */
int vpbs_mlp_infer(vpbs_mlp *net, int64_t n, const float *X, float *Y);
/* Bound of the BF16 inference's outputs: vpbs_mlp_infer (and the learned correction) clamps
 * output c to [-ybound[c], ybound[c]] (n_out positive values; NULL: no bound, the default).
 * For a network of the learned correction: the range of the training targets, so that a node
 * whose features lie outside the training data cannot receive an arbitrary correction. */
int vpbs_mlp_set_output_bound(vpbs_mlp *net, const float *ybound);
/* Loss of all n samples as one batch (mean over samples of the summed squared normalised
 * errors) and its gradient with respect to the parameters (host array of num_params floats;
 * either output may be NULL). */
int vpbs_mlp_gradient(vpbs_mlp *net, int64_t n, const float *X, const float *T, double *loss, float *grad);
/* steps Adam steps (bias-corrected) on batches of `batch` samples drawn uniformly with
 * replacement from the n samples by a hash of (seed, step, position); batch ≥ n uses all
 * samples in order. losses (host, steps values, may be NULL) receives each step's batch loss. */
int vpbs_mlp_train(vpbs_mlp *net, int64_t n, const float *X, const float *T, int32_t steps, int32_t batch, double lr,
                   double beta1, double beta2, double eps, uint64_t seed, double *losses);

/*
 * The learned correction on a context. The context's own splitting (alpha, rcut, Ns) is the
 * cheap one; alpha_conv < alpha and rcut_conv > rcut define the converged splitting on the
 * same grid (β = alpha_conv·rcut_conv = 4 is converged to ~6e-9). feature_radius: radius of
 * the strands the features describe. Every vpbs_compute_learned checks the corrected
 * velocity on ⌈sample_fraction·Np⌉ nodes (a stratified random sample, seeded by seed and the
 * call count) against the converged splitting and, when the relative L2 error exceeds tol,
 * rejects the correction and evaluates the converged splitting on every node instead. net
 * (128 inputs, 3 outputs) is borrowed and must outlive its use; it may be NULL for the
 * training helpers below. Requires filaments from vpbs_set_filaments.
 */
typedef struct vpbs_learned_report {
    double err_corrected;   /* relative L2 error of the corrected velocity vs the converged splitting over
                               the whole field, estimated from the sample: its squared errors times
                               Np / nsample over the sum of |v|^2 over every node */
    double err_uncorrected; /* the same for the cheap velocity alone */
    int64_t nsample;        /* nodes in the sample (0: no check) */
    int32_t fallback;       /* 1: correction rejected, the converged splitting was evaluated on every node */
    double total_ms;        /* wall time of the call */
} vpbs_learned_report;

int vpbs_learned_setup(vpbs_ctx *ctx, vpbs_mlp *net, double alpha_conv, double rcut_conv, double feature_radius,
                       double sample_fraction, double tol, uint64_t seed);
/* Velocity of the cheap splitting plus the network's correction, checked as above. v_out:
 * host array of 3*Np doubles or NULL (result left on the device, see vpbs_output_device_ptrs). */
int vpbs_compute_learned(vpbs_ctx *ctx, double *v_out, vpbs_learned_report *report);
/* Learned evaluations since vpbs_learned_setup (vpbs_compute_learned calls, and the
 * evaluations of vpbs_advance, which uses the learned evaluation whenever a network is set up):
 * their number, how many were rejected by the sampled check, and the largest sampled error of
 * a corrected velocity. Any pointer may be NULL. */
int vpbs_learned_stats(vpbs_ctx *ctx, int64_t *calls, int64_t *fallbacks, double *max_err);
/* Training helpers (after a vpbs_compute of the current points; pointers host or device).
 * Features of every node: X[Np][128] (6 + 6·16 used: curvature vector/α and tangent in the node's
 * frame, then for the 16 nearest strands the foot of the perpendicular on the strand's tangent
 * line (×α) and the strand's tangent, in the frame); frame[Np][9] = (t, n, b). */
int vpbs_learned_features(vpbs_ctx *ctx, float *X, float *frame);
/* Converged − cheap velocity on every node (needs VPBS_VELOCITY | VPBS_KEEP_UHAT on the last
 * vpbs_compute): dv[3*Np] (FP64) and/or T[Np][3] = its components along (t, n, b) in units of
 * Γ·alpha/4π (the network's target). */
int vpbs_learned_difference(vpbs_ctx *ctx, double *dv, float *T);

#ifdef __cplusplus
}
#endif

#endif /* VPBS_H */
