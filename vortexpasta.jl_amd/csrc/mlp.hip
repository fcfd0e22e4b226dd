// Small multilayer perceptron of the learned Biot–Savart correction: training on the GPU
// (forward, backward, Adam; FP32) and inference on BF16 MFMA with FP32 accumulation.
// It has no counterpart in the reference.
//
// Network: x (n_in) → [Linear → SiLU] × n_layers (width H) → Linear → y (n_out).
// Parameters, FP32, one flat array: for layer l = 0..n_layers the weights W_l
// [out_l][in_l] (row-major) followed by the bias b_l [out_l], with in_0 = n_in,
// out_{n_layers} = n_out and H otherwise. Inputs are normalised on the fly,
// x' = (x − in_shift)·in_scale, and the network predicts y' = y / out_scale; the
// loss of a batch of B samples is (1/B) Σ_b Σ_c (y'_bc − t_bc/out_scale_c)².
//
// Training kernels (FP32, deterministic: fixed-order reductions, no atomics):
//   sgemm<TA, TB, EPI>  64×64 tiles staged in LDS, 4×4 outputs per thread, optional
//                       split over K into partial sums reduced in order
//   loss_grad           dY' and the per-block loss partial sums
//   colsum / reduce     bias gradients and split-K sums
//   adam_step           Adam with bias correction
// Inference kernel mlp_infer_bf16: persistent blocks hold every layer's weights in LDS as
// BF16; a wavefront takes 32 samples (two 16-column tiles) at a time and runs each layer as
// v_mfma_f32_16x16x32_bf16 products with the weights as the A operand (rows = output
// neurons) and the activations as the B operand (columns = samples). The accumulator of a
// 16-neuron output block then holds, per lane, 4 consecutive neurons of one sample, which is
// half of the B fragment the next layer needs (8 consecutive K values per lane); the next
// layer's weight columns are stored permuted within each 32-neuron chunk so that two output
// blocks form one B fragment without any data movement.
#include "vpbs_internal.hpp"

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <new>
#include <vector>

struct vpbs_mlp {
    int device = 0;
    std::string err;
    int n_in = 0, H = 0, nl = 0, n_out = 0;
    int64_t nparams = 0;
    std::vector<int64_t> w_off, b_off;  // per layer offsets of W_l and b_l in the flat parameter array
    hipStream_t s = nullptr;
    float *params = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    int64_t adam_t = 0;
    float *norm = nullptr;  // in_shift[n_in], in_scale[n_in], out_scale[n_out]
    float ybound[16];       // |y_c| bound of the BF16 inference (+inf: none)
    bool have_params = false;
    // BF16 inference weights (LDS image: permuted rows with padding, then FP32 biases and the
    // normalisation), rebuilt after every parameter change
    void *infer_img = nullptr;
    size_t infer_bytes = 0;
    bool infer_dirty = true;
    // training workspace (FP32)
    float *ws = nullptr;
    int64_t ws_floats = 0;
    int32_t *idx = nullptr;
    int64_t cap_idx = 0;
    double *loss_part = nullptr;
    int64_t cap_loss = 0;
};

namespace vpbs {
namespace {

int mfail(vpbs_mlp *m, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    m->err = buf;
    return code;
}

#define MLP_HIP(m, expr)                                                                              \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return mfail((m), VPBS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                         __FILE__, __LINE__);                                                         \
    } while (0)

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ inline float sigmoidf_(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ inline float silu_(float z) { return z * sigmoidf_(z); }
// SiLU of the BF16 inference: hardware base-2 exponential and reciprocal (about 1 ulp each in
// FP32), the result being rounded to BF16 right after.
__device__ inline float silu_fast(float z) {
    return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z * 1.44269504088896340736f));
}
__device__ inline float dsilu_(float z) {
    const float s = sigmoidf_(z);
    return s * (1.0f + z * (1.0f - s));
}

// ----------------------------------------------------------------------------
// FP32 tiled matrix product: C[m][n] = Σ_k A(m, k)·B(k, n) over k in this block's K range,
// A(m, k) = TA ? A[k·lda + m] : A[m·lda + k], B(k, n) = TB ? B[n·ldb + k] : B[k·ldb + n].
enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2, EPI_DSILU = 3 };

constexpr int GT = 64, GK = 16;

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void sgemm(int M, int N, int K, const float *__restrict__ A, int lda,
                                             const float *__restrict__ B, int ldb, float *__restrict__ C, int ldc,
                                             int64_t c_split_stride, int kchunk, const float *__restrict__ bias,
                                             float *__restrict__ act, const float *__restrict__ zin) {
    __shared__ float As[GK][GT + 4];
    __shared__ float Bs[GK][GT + 4];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    C += (int64_t)blockIdx.z * c_split_stride;
    float acc[4][4] = {};
    for (int k0 = kb; k0 < ke; k0 += GK) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int e = tid + 256 * i;
            int mm, kk;
            if (TA) {
                kk = e / GT;
                mm = e % GT;
            } else {
                mm = e / GK;
                kk = e % GK;
            }
            const int gm = m0 + mm, gk = k0 + kk;
            float v = 0.f;
            if (gm < M && gk < ke) v = TA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk];
            As[kk][mm] = v;
            int nn;
            if (TB) {
                nn = e / GK;
                kk = e % GK;
            } else {
                kk = e / GT;
                nn = e % GT;
            }
            const int gn = n0 + nn, gk2 = k0 + kk;
            float w = 0.f;
            if (gn < N && gk2 < ke) w = TB ? B[(int64_t)gn * ldb + gk2] : B[(int64_t)gk2 * ldb + gn];
            Bs[kk][nn] = w;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; kk++) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                a[i] = As[kk][ty * 4 + i];
                b[i] = Bs[kk][tx * 4 + i];
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int gm = m0 + ty * 4 + i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int gn = n0 + tx * 4 + j;
            if (gn >= N) continue;
            const int64_t o = (int64_t)gm * ldc + gn;
            float v = acc[i][j];
            if (EPI == EPI_BIAS || EPI == EPI_BIAS_SILU) v += bias[gn];
            if (EPI == EPI_DSILU) v *= dsilu_(zin[o]);
            C[o] = v;
            if (EPI == EPI_BIAS_SILU) act[o] = silu_(v);
        }
    }
}

// out[i] = Σ_z part[z·stride + i], z in order
__global__ void reduce_splits(int64_t n, int nz, int64_t stride, const float *__restrict__ part,
                              float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < nz; z++) s += part[z * stride + i];
    out[i] = s;
}

// out[n] = Σ_b X[b·ld + n] (one block per column; strided partial sums, then a fixed tree)
__global__ __launch_bounds__(256) void colsum(int B, int N, const float *__restrict__ X, int ld,
                                              float *__restrict__ out) {
    __shared__ float sm[256];
    const int n = blockIdx.x;
    float s = 0.f;
    for (int b = threadIdx.x; b < B; b += 256) s += X[(int64_t)b * ld + n];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = sm[0];
}

// Batch gather with input normalisation: xb[b][k] = (X[i_b][k] − shift[k])·scale[k],
// tb[b][c] = T[i_b][c] / out_scale[c]; i_b = idx[b] (or b when idx is null).
__global__ void gather_batch(int B, int n_in, int n_out, const int32_t *__restrict__ idx,
                             const float *__restrict__ X, const float *__restrict__ T,
                             const float *__restrict__ norm, float *__restrict__ xb, float *__restrict__ tb) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nx = (int64_t)B * n_in;
    if (e < nx) {
        const int b = (int)(e / n_in), k = (int)(e % n_in);
        const int64_t i = idx ? idx[b] : b;
        xb[e] = (X[i * n_in + k] - norm[k]) * norm[n_in + k];
    } else if (T && e < nx + (int64_t)B * n_out) {
        const int64_t f = e - nx;
        const int b = (int)(f / n_out), c = (int)(f % n_out);
        const int64_t i = idx ? idx[b] : b;
        tb[f] = T[i * n_out + c] / norm[2 * n_in + c];
    }
}

__device__ inline uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Batch indices drawn uniformly (with replacement) from [0, n): a hash of (seed, step, b).
__global__ void draw_indices(int B, int64_t n, uint64_t seed, int64_t step, int32_t *__restrict__ idx) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const uint64_t h = splitmix(seed ^ splitmix((uint64_t)step * 0x100000001B3ull + (uint64_t)b));
    idx[b] = (int32_t)(h % (uint64_t)n);
}

// dY'[b][c] = 2(y' − t')/B; per-block partial sums of (y' − t')² (fixed tree order).
__global__ __launch_bounds__(256) void loss_grad(int B, int n_out, const float *__restrict__ y,
                                                 const float *__restrict__ t, float *__restrict__ dy,
                                                 double *__restrict__ part) {
    __shared__ double sm[256];
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double l = 0.0;
    if (e < (int64_t)B * n_out) {
        const float d = y[e] - t[e];
        dy[e] = 2.0f * d / (float)B;
        l = (double)d * (double)d;
    }
    sm[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void adam_step(int64_t n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    // c1 = 1/(1 − β1^t), c2 = 1/(1 − β2^t)
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// ----------------------------------------------------------------------------
// BF16 inference

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t to_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
    return (uint16_t)(u >> 16);
}

__device__ inline uint32_t pack_bf16(float lo, float hi) {
    return (uint32_t)to_bf16(lo) | ((uint32_t)to_bf16(hi) << 16);
}

__device__ inline bf16x8 as_bf16x8(uint4 u) { return __builtin_bit_cast(bf16x8, u); }

// LDS image of the inference weights. Rows of W_l are padded to in + 8 BF16 values (row
// starts shift by 4 banks, so the 16 rows an A-fragment load touches hit distinct banks).
// Columns of W_l (l ≥ 1) are permuted within each 32-column chunk: position 8g + 4h + i
// holds logical input 16h + 4g + i (see the header comment). The output layer is padded to
// 16 rows. After the weights: the FP32 biases (H per hidden layer, 16 for the output), the
// normalisation (in_shift, in_scale: n_in each; out_scale: 16) and the output bound (16).
struct InferLayout {
    int n_in, H, nl, n_out;
    size_t w_off[8];  // BF16 offsets of each layer's rows
    size_t b_off;     // byte offset of the FP32 block
    size_t bytes;
};

inline InferLayout infer_layout(int n_in, int H, int nl, int n_out) {
    InferLayout L{};
    L.n_in = n_in;
    L.H = H;
    L.nl = nl;
    L.n_out = n_out;
    size_t off = 0;
    for (int l = 0; l <= nl; l++) {
        const int in = l == 0 ? n_in : H, out = l == nl ? 16 : H;
        L.w_off[l] = off;
        off += (size_t)out * (in + 8);
    }
    L.b_off = ((off * 2 + 15) / 16) * 16;
    const size_t nf = (size_t)nl * H + 16 + 2 * (size_t)n_in + 16 + 16;
    L.bytes = L.b_off + nf * 4;
    return L;
}

constexpr int INFER_WAVES = 8;  // wavefronts per block
constexpr int INFER_TILES = 2;  // 16-sample tiles per wavefront step

// Y[i][c] = out_scale[c]·net(x'_i)[c] for c < n_out, clamped to ±bound[c] (FP32 out), X rows of
// n_in floats.
template <int NIN, int H, int NL>
__global__ __launch_bounds__(64 * INFER_WAVES) void mlp_infer_bf16(int64_t n, int n_out, const float *__restrict__ X,
                                                                    const uint4 *__restrict__ img, int img_vec,
                                                                    float *__restrict__ Y) {
    extern __shared__ uint4 lds4[];
    // copy the image into LDS (16-byte vectors)
    for (int e = threadIdx.x; e < img_vec; e += blockDim.x) lds4[e] = img[e];
    __syncthreads();
    const uint16_t *W = (const uint16_t *)lds4;
    constexpr int RIN0 = NIN + 8, RH = H + 8;
    // layer row bases (BF16 units), as in infer_layout
    constexpr size_t W0 = 0, W1 = (size_t)H * RIN0;
    constexpr size_t WBYTES = 2 * (W1 + (size_t)(NL - 1) * H * RH + (size_t)16 * RH);
    constexpr size_t BOFF = ((WBYTES + 15) / 16) * 16;
    const float *fb = (const float *)((const char *)lds4 + BOFF);
    const float *bias_out = fb + NL * H;
    const float *shift = bias_out + 16, *scale = shift + NIN, *oscale = scale + NIN, *obound = oscale + 16;

    const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    const int64_t wave = (int64_t)blockIdx.x * INFER_WAVES + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * INFER_WAVES;
    constexpr int KC0 = NIN / 32, KCH = H / 32;

    for (int64_t base = wave * 16 * INFER_TILES; base < n; base += nwaves * 16 * INFER_TILES) {
        // B fragments of the input: sample base + 16t + col, features 32c + 8g .. + 7
        bf16x8 act[INFER_TILES][(KC0 > KCH ? KC0 : KCH)];
#pragma unroll
        for (int t = 0; t < INFER_TILES; t++) {
            const int64_t i = base + 16 * t + col;
#pragma unroll
            for (int c = 0; c < KC0; c++) {
                const int k = 32 * c + 8 * g;
                float x[8];
                if (i < n) {
                    const float4 *p = (const float4 *)(X + i * NIN + k);
                    const float4 u = p[0], v = p[1];
                    x[0] = u.x, x[1] = u.y, x[2] = u.z, x[3] = u.w, x[4] = v.x, x[5] = v.y, x[6] = v.z, x[7] = v.w;
                } else {
#pragma unroll
                    for (int q = 0; q < 8; q++) x[q] = shift[k + q];  // normalises to 0
                }
                uint4 pk;
                pk.x = pack_bf16((x[0] - shift[k]) * scale[k], (x[1] - shift[k + 1]) * scale[k + 1]);
                pk.y = pack_bf16((x[2] - shift[k + 2]) * scale[k + 2], (x[3] - shift[k + 3]) * scale[k + 3]);
                pk.z = pack_bf16((x[4] - shift[k + 4]) * scale[k + 4], (x[5] - shift[k + 5]) * scale[k + 5]);
                pk.w = pack_bf16((x[6] - shift[k + 6]) * scale[k + 6], (x[7] - shift[k + 7]) * scale[k + 7]);
                act[t][c] = as_bf16x8(pk);
            }
        }
        // hidden layers
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const int kc = l == 0 ? KC0 : KCH;
            const int rs = l == 0 ? RIN0 : RH;
            const uint16_t *Wl = W + (l == 0 ? W0 : W1 + (size_t)(l - 1) * H * RH);
            const float *bl = fb + l * H;
            f32x4 acc[INFER_TILES][H / 16];
#pragma unroll
            for (int mb = 0; mb < H / 16; mb++) {
                f32x4 b0;
#pragma unroll
                for (int v = 0; v < 4; v++) b0[v] = bl[16 * mb + 4 * g + v];
#pragma unroll
                for (int t = 0; t < INFER_TILES; t++) acc[t][mb] = b0;
#pragma unroll
                for (int c = 0; c < kc; c++) {
                    const bf16x8 a = as_bf16x8(*(const uint4 *)(Wl + (size_t)(16 * mb + col) * rs + 32 * c + 8 * g));
#pragma unroll
                    for (int t = 0; t < INFER_TILES; t++)
                        acc[t][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, act[t][c], acc[t][mb], 0, 0, 0);
                }
            }
            // SiLU; output blocks 2c and 2c + 1 form the next layer's K chunk c
#pragma unroll
            for (int t = 0; t < INFER_TILES; t++)
#pragma unroll
                for (int c = 0; c < KCH; c++) {
                    const f32x4 lo = acc[t][2 * c], hi = acc[t][2 * c + 1];
                    uint4 pk;
                    pk.x = pack_bf16(silu_fast(lo[0]), silu_fast(lo[1]));
                    pk.y = pack_bf16(silu_fast(lo[2]), silu_fast(lo[3]));
                    pk.z = pack_bf16(silu_fast(hi[0]), silu_fast(hi[1]));
                    pk.w = pack_bf16(silu_fast(hi[2]), silu_fast(hi[3]));
                    act[t][c] = as_bf16x8(pk);
                }
        }
        // output layer (16 padded rows): lane (g, v) holds output 4g + v of sample col
        const uint16_t *Wo = W + W1 + (size_t)(NL - 1) * H * RH;
        f32x4 out[INFER_TILES];
        f32x4 bo;
#pragma unroll
        for (int v = 0; v < 4; v++) bo[v] = bias_out[4 * g + v];
#pragma unroll
        for (int t = 0; t < INFER_TILES; t++) out[t] = bo;
#pragma unroll
        for (int c = 0; c < KCH; c++) {
            const bf16x8 a = as_bf16x8(*(const uint4 *)(Wo + (size_t)col * RH + 32 * c + 8 * g));
#pragma unroll
            for (int t = 0; t < INFER_TILES; t++) out[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, act[t][c], out[t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < INFER_TILES; t++) {
            const int64_t i = base + 16 * t + col;
            if (i >= n) continue;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int o = 4 * g + v;
                if (o < n_out) Y[i * n_out + o] = fminf(fmaxf(out[t][v] * oscale[o], -obound[o]), obound[o]);
            }
        }
    }
}

// ----------------------------------------------------------------------------
// host side

int64_t layer_in(const vpbs_mlp *m, int l) { return l == 0 ? m->n_in : m->H; }
int64_t layer_out(const vpbs_mlp *m, int l) { return l == m->nl ? m->n_out : m->H; }

// Workspace for a batch of B samples: xb [B][n_in], tb [B][n_out], per hidden layer Z and A
// [B][H], y [B][n_out], dy [B][n_out], two dA/dZ buffers [B][H], split-K partials.
struct Work {
    float *xb, *tb, *y, *dy, *d0, *d1, *part;
    std::vector<float *> Z, A;
};
constexpr int MAX_SPLIT = 64;

int64_t work_floats(const vpbs_mlp *m, int64_t B) {
    const int64_t wmax = std::max<int64_t>((int64_t)m->H * std::max(m->n_in, m->H), (int64_t)m->n_out * m->H);
    return B * (m->n_in + 2 * m->n_out + 2 * (int64_t)m->H * m->nl + 2 * (int64_t)m->H) + MAX_SPLIT * wmax;
}

int ensure_work(vpbs_mlp *m, int64_t B, Work &w) {
    const int64_t need = work_floats(m, B);
    if (need > m->ws_floats) {
        if (m->ws) MLP_HIP(m, hipFree(m->ws));
        m->ws = nullptr;
        MLP_HIP(m, hipMalloc((void **)&m->ws, sizeof(float) * (size_t)need));
        m->ws_floats = need;
    }
    if (B > m->cap_idx) {
        if (m->idx) MLP_HIP(m, hipFree(m->idx));
        m->idx = nullptr;
        MLP_HIP(m, hipMalloc((void **)&m->idx, sizeof(int32_t) * (size_t)B));
        m->cap_idx = B;
    }
    const int64_t nlb = (B * m->n_out + 255) / 256;
    if (nlb > m->cap_loss) {
        if (m->loss_part) MLP_HIP(m, hipFree(m->loss_part));
        m->loss_part = nullptr;
        MLP_HIP(m, hipMalloc((void **)&m->loss_part, sizeof(double) * (size_t)nlb));
        m->cap_loss = nlb;
    }
    float *p = m->ws;
    w.xb = p, p += B * m->n_in;
    w.tb = p, p += B * m->n_out;
    w.y = p, p += B * m->n_out;
    w.dy = p, p += B * m->n_out;
    w.Z.resize(m->nl);
    w.A.resize(m->nl);
    for (int l = 0; l < m->nl; l++) {
        w.Z[l] = p, p += B * m->H;
        w.A[l] = p, p += B * m->H;
    }
    w.d0 = p, p += B * m->H;
    w.d1 = p, p += B * m->H;
    w.part = p;
    return VPBS_OK;
}

template <bool TA, bool TB, int EPI>
void gemm(hipStream_t s, int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C, int ldc,
          int nsplit, int64_t split_stride, const float *bias = nullptr, float *act = nullptr,
          const float *zin = nullptr) {
    int kchunk = (K + nsplit - 1) / nsplit;
    kchunk = ((kchunk + GK - 1) / GK) * GK;
    nsplit = (K + kchunk - 1) / kchunk;
    dim3 grid(nblk(N, GT), nblk(M, GT), nsplit);
    sgemm<TA, TB, EPI><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, split_stride, kchunk, bias, act, zin);
}

// Forward pass of a normalised batch xb (B samples): Z, A of the hidden layers and y'.
void forward(vpbs_mlp *m, int64_t B, const Work &w) {
    const float *in = w.xb;
    int in_dim = m->n_in;
    for (int l = 0; l < m->nl; l++) {
        const float *W = m->params + m->w_off[l], *b = m->params + m->b_off[l];
        gemm<false, true, EPI_BIAS_SILU>(m->s, (int)B, m->H, in_dim, in, in_dim, W, in_dim, w.Z[l], m->H, 1, 0, b,
                                         w.A[l]);
        in = w.A[l];
        in_dim = m->H;
    }
    const int L = m->nl;
    gemm<false, true, EPI_BIAS>(m->s, (int)B, m->n_out, in_dim, in, in_dim, m->params + m->w_off[L], in_dim, w.y,
                                m->n_out, 1, 0, m->params + m->b_off[L]);
}

// Loss and gradient of a normalised batch (after forward): gradient into m->grad, loss
// partials into m->loss_part (returned as the number of partials).
int64_t backward(vpbs_mlp *m, int64_t B, const Work &w) {
    const int64_t nlb = (B * m->n_out + 255) / 256;
    loss_grad<<<(unsigned)nlb, 256, 0, m->s>>>((int)B, m->n_out, w.y, w.tb, w.dy, m->loss_part);
    const int nsplit = (int)std::min<int64_t>(MAX_SPLIT, std::max<int64_t>(1, B / 512));
    const float *dz = w.dy;  // gradient w.r.t. the current layer's pre-activation
    int out_dim = m->n_out;
    float *dbuf[2] = {w.d0, w.d1};
    for (int l = m->nl; l >= 0; l--) {
        const int in_dim = (int)layer_in(m, l);
        const float *ain = l == 0 ? w.xb : w.A[l - 1];
        const int64_t wsz = (int64_t)out_dim * in_dim;
        // dW_l = dZ^T · A_in (split over the batch, partial sums reduced in order)
        gemm<true, false, EPI_STORE>(m->s, out_dim, in_dim, (int)B, dz, out_dim, ain, in_dim, w.part, in_dim, nsplit,
                                     wsz);
        int kchunk = (int)((B + nsplit - 1) / nsplit);
        kchunk = ((kchunk + GK - 1) / GK) * GK;
        const int nz = (int)((B + kchunk - 1) / kchunk);
        reduce_splits<<<nblk(wsz, 256), 256, 0, m->s>>>(wsz, nz, wsz, w.part, m->grad + m->w_off[l]);
        colsum<<<out_dim, 256, 0, m->s>>>((int)B, out_dim, dz, out_dim, m->grad + m->b_off[l]);
        if (l == 0) break;
        // dA_{l-1} = dZ · W_l, then dZ_{l-1} = dA ⊙ silu'(Z_{l-1})
        float *dn = dbuf[l & 1];
        gemm<false, false, EPI_DSILU>(m->s, (int)B, in_dim, out_dim, dz, out_dim, m->params + m->w_off[l], in_dim, dn,
                                      in_dim, 1, 0, nullptr, nullptr, w.Z[l - 1]);
        dz = dn;
        out_dim = in_dim;
    }
    return nlb;
}

int sum_loss(vpbs_mlp *m, int64_t nlb, double *loss) {
    std::vector<double> h((size_t)nlb);
    MLP_HIP(m, hipMemcpyAsync(h.data(), m->loss_part, sizeof(double) * (size_t)nlb, hipMemcpyDeviceToHost, m->s));
    MLP_HIP(m, hipStreamSynchronize(m->s));
    double s = 0;
    for (double v : h) s += v;
    *loss = s;
    return VPBS_OK;
}

template <int NIN, int H, int NL>
int launch_infer(vpbs_mlp *m, int64_t n, const float *X, float *Y, hipStream_t s) {
    const int vec = (int)(m->infer_bytes / 16);
    static bool attr_set = false;
    if (!attr_set) {
        MLP_HIP(m, hipFuncSetAttribute((const void *)mlp_infer_bf16<NIN, H, NL>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)m->infer_bytes));
        attr_set = true;
    }
    int ncu = 256;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, m->device);
    const int64_t per_block = 16 * INFER_TILES * INFER_WAVES;
    const int64_t nb = std::min<int64_t>((int64_t)ncu, (n + per_block - 1) / per_block);
    mlp_infer_bf16<NIN, H, NL><<<(unsigned)std::max<int64_t>(nb, 1), 64 * INFER_WAVES, m->infer_bytes, s>>>(
        n, m->n_out, X, (const uint4 *)m->infer_img, vec, Y);
    MLP_HIP(m, hipGetLastError());
    return VPBS_OK;
}

// Rebuild the BF16 LDS image from the FP32 parameters and the normalisation.
int build_infer_image(vpbs_mlp *m) {
    if (!m->infer_dirty) return VPBS_OK;
    const InferLayout L = infer_layout(m->n_in, m->H, m->nl, m->n_out);
    std::vector<float> P((size_t)m->nparams), N((size_t)(2 * m->n_in + m->n_out));
    MLP_HIP(m, hipMemcpy(P.data(), m->params, sizeof(float) * P.size(), hipMemcpyDeviceToHost));
    MLP_HIP(m, hipMemcpy(N.data(), m->norm, sizeof(float) * N.size(), hipMemcpyDeviceToHost));
    std::vector<unsigned char> img(L.bytes, 0);
    uint16_t *w = (uint16_t *)img.data();
    for (int l = 0; l <= m->nl; l++) {
        const int in = (int)layer_in(m, l), out = (int)layer_out(m, l), rs = in + 8;
        const float *W = P.data() + m->w_off[l];
        for (int o = 0; o < out; o++)
            for (int p = 0; p < in; p++) {
                int k = p;  // logical input column stored at position p
                if (l > 0) {
                    const int q = p % 32, g = q / 8, h = (q / 4) % 2, i = q % 4;
                    k = (p - q) + 16 * h + 4 * g + i;
                }
                w[L.w_off[l] + (size_t)o * rs + p] = to_bf16(W[(size_t)o * in + k]);
            }
    }
    float *fb = (float *)(img.data() + L.b_off);
    for (int l = 0; l < m->nl; l++)
        for (int o = 0; o < m->H; o++) fb[l * m->H + o] = P[m->b_off[l] + o];
    float *bo = fb + m->nl * m->H;
    for (int o = 0; o < m->n_out; o++) bo[o] = P[m->b_off[m->nl] + o];
    float *sh = bo + 16, *sc = sh + m->n_in, *os = sc + m->n_in;
    for (int k = 0; k < m->n_in; k++) {
        sh[k] = N[k];
        sc[k] = N[m->n_in + k];
    }
    for (int o = 0; o < m->n_out; o++) os[o] = N[2 * m->n_in + o];
    float *ob = os + 16;
    for (int o = 0; o < 16; o++) ob[o] = m->ybound[o];
    if (m->infer_img && m->infer_bytes != L.bytes) {
        MLP_HIP(m, hipFree(m->infer_img));
        m->infer_img = nullptr;
    }
    if (!m->infer_img) MLP_HIP(m, hipMalloc(&m->infer_img, L.bytes));
    m->infer_bytes = L.bytes;
    MLP_HIP(m, hipMemcpy(m->infer_img, img.data(), L.bytes, hipMemcpyHostToDevice));
    m->infer_dirty = false;
    return VPBS_OK;
}

}  // namespace

// Inference of the learned correction on samples X (device, n rows of n_in floats) into Y
// (device, n rows of n_out floats), on stream s.
int mlp_infer(vpbs_mlp *m, int64_t n, const float *X, float *Y, hipStream_t s) {
    if (!m->have_params) return mfail(m, VPBS_ERR_STATE, "network parameters not set");
    MLP_HIP(m, hipSetDevice(m->device));
    if (n <= 0) return VPBS_OK;
    int rc = build_infer_image(m);
    if (rc != VPBS_OK) return rc;
    if (m->n_in == 128 && m->H == 128 && m->nl == 3) return launch_infer<128, 128, 3>(m, n, X, Y, s);
    if (m->n_in == 128 && m->H == 128 && m->nl == 2) return launch_infer<128, 128, 2>(m, n, X, Y, s);
    if (m->n_in == 32 && m->H == 32 && m->nl == 2) return launch_infer<32, 32, 2>(m, n, X, Y, s);
    return mfail(m, VPBS_ERR_INVALID, "BF16 inference is compiled for (n_in, H, layers) = (128, 128, 3), "
                 "(128, 128, 2) and (32, 32, 2)");
}

const char *mlp_error(const vpbs_mlp *m) { return m->err.c_str(); }
int mlp_n_in(const vpbs_mlp *m) { return m->n_in; }
int mlp_n_out(const vpbs_mlp *m) { return m->n_out; }

}  // namespace vpbs

using namespace vpbs;

extern "C" {

int vpbs_mlp_create(vpbs_mlp **out, int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out,
                    int32_t device) {
    if (!out) return VPBS_ERR_INVALID;
    *out = nullptr;
    if (n_in < 1 || n_in % 32 || n_hidden < 32 || n_hidden % 32 || n_layers < 1 || n_layers > 7 || n_out < 1 ||
        n_out > 16)
        return VPBS_ERR_INVALID;
    vpbs_mlp *m = new (std::nothrow) vpbs_mlp();
    if (!m) return VPBS_ERR_HIP;
    m->device = device;
    m->n_in = n_in;
    m->H = n_hidden;
    m->nl = n_layers;
    m->n_out = n_out;
    int64_t off = 0;
    for (int l = 0; l <= n_layers; l++) {
        m->w_off.push_back(off);
        off += layer_out(m, l) * layer_in(m, l);
        m->b_off.push_back(off);
        off += layer_out(m, l);
    }
    m->nparams = off;
    for (int o = 0; o < 16; o++) m->ybound[o] = INFINITY;
    bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->params, sizeof(float) * (size_t)off) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->grad, sizeof(float) * (size_t)off) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->adam_m, sizeof(float) * (size_t)off) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->adam_v, sizeof(float) * (size_t)off) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->norm, sizeof(float) * (size_t)(2 * n_in + n_out)) == hipSuccess;
    if (!ok) {
        vpbs_mlp_destroy(m);
        return VPBS_ERR_HIP;
    }
    *out = m;
    return VPBS_OK;
}

void vpbs_mlp_destroy(vpbs_mlp *m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->s) (void)hipStreamSynchronize(m->s);
    void *bufs[] = {m->params, m->grad, m->adam_m, m->adam_v, m->norm, m->infer_img, m->ws, m->idx, m->loss_part};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
}

const char *vpbs_mlp_last_error(const vpbs_mlp *m) { return m ? m->err.c_str() : ""; }

int64_t vpbs_mlp_num_params(const vpbs_mlp *m) { return m ? m->nparams : -1; }

int vpbs_mlp_set_params(vpbs_mlp *m, const float *params, const float *in_shift, const float *in_scale,
                        const float *out_scale) {
    if (!m) return VPBS_ERR_INVALID;
    if (!params || !in_shift || !in_scale || !out_scale) return mfail(m, VPBS_ERR_INVALID, "null argument");
    MLP_HIP(m, hipSetDevice(m->device));
    MLP_HIP(m, hipMemcpy(m->params, params, sizeof(float) * (size_t)m->nparams, hipMemcpyHostToDevice));
    MLP_HIP(m, hipMemcpy(m->norm, in_shift, sizeof(float) * m->n_in, hipMemcpyHostToDevice));
    MLP_HIP(m, hipMemcpy(m->norm + m->n_in, in_scale, sizeof(float) * m->n_in, hipMemcpyHostToDevice));
    MLP_HIP(m, hipMemcpy(m->norm + 2 * m->n_in, out_scale, sizeof(float) * m->n_out, hipMemcpyHostToDevice));
    MLP_HIP(m, hipMemset(m->adam_m, 0, sizeof(float) * (size_t)m->nparams));
    MLP_HIP(m, hipMemset(m->adam_v, 0, sizeof(float) * (size_t)m->nparams));
    m->adam_t = 0;
    m->have_params = true;
    m->infer_dirty = true;
    return VPBS_OK;
}

int vpbs_mlp_get_params(vpbs_mlp *m, float *params, float *in_shift, float *in_scale, float *out_scale) {
    if (!m) return VPBS_ERR_INVALID;
    if (!m->have_params) return mfail(m, VPBS_ERR_STATE, "network parameters not set");
    MLP_HIP(m, hipSetDevice(m->device));
    MLP_HIP(m, hipStreamSynchronize(m->s));
    if (params) MLP_HIP(m, hipMemcpy(params, m->params, sizeof(float) * (size_t)m->nparams, hipMemcpyDeviceToHost));
    if (in_shift) MLP_HIP(m, hipMemcpy(in_shift, m->norm, sizeof(float) * m->n_in, hipMemcpyDeviceToHost));
    if (in_scale) MLP_HIP(m, hipMemcpy(in_scale, m->norm + m->n_in, sizeof(float) * m->n_in, hipMemcpyDeviceToHost));
    if (out_scale)
        MLP_HIP(m, hipMemcpy(out_scale, m->norm + 2 * m->n_in, sizeof(float) * m->n_out, hipMemcpyDeviceToHost));
    return VPBS_OK;
}

int vpbs_mlp_set_output_bound(vpbs_mlp *m, const float *ybound) {
    if (!m) return VPBS_ERR_INVALID;
    for (int o = 0; o < m->n_out; o++) {
        const float b = ybound ? ybound[o] : INFINITY;
        if (!(b > 0)) return mfail(m, VPBS_ERR_INVALID, "output bound %d must be positive (got %g)", o, (double)b);
        m->ybound[o] = b;
    }
    m->infer_dirty = true;
    return VPBS_OK;
}

int vpbs_mlp_forward(vpbs_mlp *m, int64_t n, const float *X, float *Y) {
    if (!m) return VPBS_ERR_INVALID;
    if (!m->have_params) return mfail(m, VPBS_ERR_STATE, "network parameters not set");
    if (n < 0 || (n > 0 && (!X || !Y))) return mfail(m, VPBS_ERR_INVALID, "invalid samples");
    if (n > (1 << 24)) return mfail(m, VPBS_ERR_INVALID, "at most 2^24 samples per call");
    MLP_HIP(m, hipSetDevice(m->device));
    if (n == 0) return VPBS_OK;
    Work w;
    int rc = ensure_work(m, n, w);
    if (rc != VPBS_OK) return rc;
    const int64_t ne = n * m->n_in;
    gather_batch<<<nblk(ne, 256), 256, 0, m->s>>>((int)n, m->n_in, m->n_out, nullptr, X, nullptr, m->norm, w.xb,
                                                  nullptr);
    forward(m, n, w);
    // y = y'·out_scale
    std::vector<float> os((size_t)m->n_out);
    MLP_HIP(m, hipMemcpyAsync(os.data(), m->norm + 2 * m->n_in, sizeof(float) * m->n_out, hipMemcpyDeviceToHost, m->s));
    MLP_HIP(m, hipStreamSynchronize(m->s));
    std::vector<float> y((size_t)(n * m->n_out));
    MLP_HIP(m, hipMemcpy(y.data(), w.y, sizeof(float) * y.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++)
        for (int c = 0; c < m->n_out; c++) y[(size_t)(i * m->n_out + c)] *= os[(size_t)c];
    MLP_HIP(m, hipMemcpy(Y, y.data(), sizeof(float) * y.size(), hipMemcpyHostToDevice));
    return VPBS_OK;
}

int vpbs_mlp_infer(vpbs_mlp *m, int64_t n, const float *X, float *Y) {
    if (!m) return VPBS_ERR_INVALID;
    if (n < 0 || (n > 0 && (!X || !Y))) return mfail(m, VPBS_ERR_INVALID, "invalid samples");
    int rc = mlp_infer(m, n, X, Y, m->s);
    if (rc != VPBS_OK) return rc;
    MLP_HIP(m, hipStreamSynchronize(m->s));
    return VPBS_OK;
}

int vpbs_mlp_gradient(vpbs_mlp *m, int64_t n, const float *X, const float *T, double *loss, float *grad) {
    if (!m) return VPBS_ERR_INVALID;
    if (!m->have_params) return mfail(m, VPBS_ERR_STATE, "network parameters not set");
    if (n < 1 || !X || !T) return mfail(m, VPBS_ERR_INVALID, "invalid samples");
    if (n > (1 << 24)) return mfail(m, VPBS_ERR_INVALID, "at most 2^24 samples per call");
    MLP_HIP(m, hipSetDevice(m->device));
    Work w;
    int rc = ensure_work(m, n, w);
    if (rc != VPBS_OK) return rc;
    const int64_t ne = n * (m->n_in + m->n_out);
    gather_batch<<<nblk(ne, 256), 256, 0, m->s>>>((int)n, m->n_in, m->n_out, nullptr, X, T, m->norm, w.xb, w.tb);
    forward(m, n, w);
    const int64_t nlb = backward(m, n, w);
    MLP_HIP(m, hipGetLastError());
    double l = 0;
    rc = sum_loss(m, nlb, &l);
    if (rc != VPBS_OK) return rc;
    if (loss) *loss = l / (double)n;
    if (grad) MLP_HIP(m, hipMemcpy(grad, m->grad, sizeof(float) * (size_t)m->nparams, hipMemcpyDeviceToHost));
    return VPBS_OK;
}

int vpbs_mlp_train(vpbs_mlp *m, int64_t n, const float *X, const float *T, int32_t steps, int32_t batch, double lr,
                   double beta1, double beta2, double eps, uint64_t seed, double *losses) {
    if (!m) return VPBS_ERR_INVALID;
    if (!m->have_params) return mfail(m, VPBS_ERR_STATE, "network parameters not set");
    if (n < 1 || !X || !T || steps < 0 || batch < 1) return mfail(m, VPBS_ERR_INVALID, "invalid arguments");
    if (n > INT32_MAX) return mfail(m, VPBS_ERR_INVALID, "at most 2^31 - 1 samples");
    if (batch > (1 << 20)) return mfail(m, VPBS_ERR_INVALID, "batch above 2^20 samples");
    if (!(lr > 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
        return mfail(m, VPBS_ERR_INVALID, "invalid Adam parameters");
    MLP_HIP(m, hipSetDevice(m->device));
    const bool full = batch >= n;  // whole data set in order, every step
    const int64_t B = full ? n : batch;
    Work w;
    int rc = ensure_work(m, B, w);
    if (rc != VPBS_OK) return rc;
    std::vector<double> part;
    for (int32_t st = 0; st < steps; st++) {
        if (!full) draw_indices<<<nblk(B, 256), 256, 0, m->s>>>((int)B, n, seed, m->adam_t, m->idx);
        const int64_t ne = B * (m->n_in + m->n_out);
        gather_batch<<<nblk(ne, 256), 256, 0, m->s>>>((int)B, m->n_in, m->n_out, full ? nullptr : m->idx, X, T,
                                                      m->norm, w.xb, w.tb);
        forward(m, B, w);
        const int64_t nlb = backward(m, B, w);
        m->adam_t++;
        const double c1 = 1.0 / (1.0 - std::pow(beta1, (double)m->adam_t));
        const double c2 = 1.0 / (1.0 - std::pow(beta2, (double)m->adam_t));
        adam_step<<<nblk(m->nparams, 256), 256, 0, m->s>>>(m->nparams, m->params, m->grad, m->adam_m, m->adam_v,
                                                           (float)lr, (float)beta1, (float)beta2, (float)eps,
                                                           (float)c1, (float)c2);
        MLP_HIP(m, hipGetLastError());
        if (losses) {
            double l = 0;
            rc = sum_loss(m, nlb, &l);
            if (rc != VPBS_OK) return rc;
            losses[st] = l / (double)B;
        }
    }
    MLP_HIP(m, hipStreamSynchronize(m->s));
    m->infer_dirty = true;
    return VPBS_OK;
}

}  // extern "C"
