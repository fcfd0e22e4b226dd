// Spectral part of the long-range NUFFT on gfx950: hand-written FFTs fused with
// the pencil reorderings and with the Fourier-space kernel.
//
// The real spreading grid holds pairs of y-rows as one complex row (pair p:
// value (x, y, z) at ((z·G1/2 + y/2)·G0 + x)·2 + (y & 1)), so one complex
// transform along x serves two real rows. Five kernels, each a single pass over
// HBM with the transform done in LDS:
//
//   spec_x_forward  pair rows → FFT_x → A_k = (Z_k + conj Z_-k)/2, B_k = (Z_k - conj Z_-k)/2i
//                   for the retained kx ∈ [0, nkx) → x1[c][z][y][kx]
//   spec_y_forward  x1 tiles [y][kx-block] → FFT_y → retained ky → x2[c][jy][z][kx]
//   spec_z_fourier  x2 tiles [z][kx-block] of the 3 components → FFT_z → deconvolution,
//                   Γ/V, Green function 1/k², Ewald factor e^{-k²/4α²}, curl i k ×
//                   (velocity) and/or nothing (streamfunction) on the retained modes,
//                   optionally the unsmoothed field into uhat → inverse FFT_z → x2 (velocity)
//                   and x2b (streamfunction), in z space
//   spec_y_inverse  x2 → zero-padded ky rows → inverse FFT_y → x1
//   spec_x_inverse  x1 → pair rows Z_k = A_k + i B_k (Hermitian half rebuilt, DC real)
//                   → inverse FFT_x → real grid
//
// These restate the transforms of the reference's NUFFT back end
// (longrange.jl:488-495 type 1, 360-378 type 2, with the Fourier-space operators
// of longrange.jl:102-156 and the Ewald factor of cache_common.jl:64-76,
// restricted to the modes kept by longrange.jl:556-583).
//
// Transforms: Stockham autosort (decimation in time) over mixed radices
// {16, 8, 4, 2, 3, 5} (every grid size is 32·2^a·3^b·5^c), done on a tile of B
// columns in LDS, S[row·B + b]: consecutive lanes take consecutive columns, so
// LDS accesses run along rows. Each pass reads its R inputs per butterfly into
// registers, synchronises, applies twiddles from an LDS table e^{-2πin/L}, does
// the R-point DFT and writes the outputs back in place. Tiles hold at most 6144
// (FP32) or 3072 (FP64) complex values: B = that / L, a power of two.

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "vpbs_internal.hpp"
#include "device_util.hpp"

namespace vpbs {
namespace {

// Kernel kinds: x kernels (rows along the contiguous axis), y kernels, z kernels
// (three component tiles). Threads per block and complex values per tile of each;
// the tiles are sized for several blocks per CU (x: 4-5, y: 2, z: 2).
enum { KIND_X = 0, KIND_Y = 1, KIND_Z = 2 };
constexpr int kind_nt(int kind) { return kind == KIND_Y ? 512 : 256; }
template <typename CT>
constexpr int kind_tile(int kind) {
    return (sizeof(CT) == 4 ? 6144 : 3072) / (kind == KIND_Y ? 1 : 2);
}
constexpr int NTX = kind_nt(KIND_X), NTY = kind_nt(KIND_Y), NTZ = kind_nt(KIND_Z);
// LDS budget of the z kernels: two blocks per CU
constexpr size_t SPEC_LDS_Z = 80 * 1024;

template <typename CT>
struct C2 {
    CT re, im;
};

// two consecutive complex values, moved with 16-byte vector accesses (one in FP32,
// two in FP64) by load_pair / store_pair; p must be 16-byte aligned
template <typename CT>
struct C2x2 {
    C2<CT> a, b;
};

// Pair transfers pay off in FP32 (one 16-byte access instead of two 8-byte ones);
// FP64 values are 16 bytes each already.
template <typename CT>
constexpr bool use_pairs() {
    return sizeof(CT) == 4;
}

template <typename CT>
__device__ __forceinline__ C2x2<CT> load_pair(const C2<CT> *p) {
    if constexpr (sizeof(CT) == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(p);
        return {{v.x, v.y}, {v.z, v.w}};
    } else {
        const double2 u = reinterpret_cast<const double2 *>(p)[0], v = reinterpret_cast<const double2 *>(p)[1];
        return {{u.x, u.y}, {v.x, v.y}};
    }
}

template <typename CT>
__device__ __forceinline__ void store_pair(C2<CT> *p, const C2x2<CT> &v) {
    if constexpr (sizeof(CT) == 4) {
        *reinterpret_cast<float4 *>(p) = float4{v.a.re, v.a.im, v.b.re, v.b.im};
    } else {
        reinterpret_cast<double2 *>(p)[0] = double2{v.a.re, v.a.im};
        reinterpret_cast<double2 *>(p)[1] = double2{v.b.re, v.b.im};
    }
}

template <typename CT>
__device__ __forceinline__ C2<CT> cadd(C2<CT> a, C2<CT> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename CT>
__device__ __forceinline__ C2<CT> csub(C2<CT> a, C2<CT> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename CT>
__device__ __forceinline__ C2<CT> cmul(C2<CT> a, C2<CT> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// multiply by S·i (S = ±1)
template <int S, typename CT>
__device__ __forceinline__ C2<CT> mul_si(C2<CT> a) {
    return S > 0 ? C2<CT>{-a.im, a.re} : C2<CT>{a.im, -a.re};
}
template <typename CT>
__device__ __forceinline__ C2<CT> cscale(C2<CT> a, CT s) {
    return {a.re * s, a.im * s};
}

// Dynamic LDS of every kernel here: the tile(s), then the twiddle table. The
// transform routines address it through this symbol with element offsets, so
// that their accesses compile to LDS instructions also out of line.
extern __shared__ __align__(16) unsigned char spec_smem[];

template <typename CT>
__device__ __forceinline__ C2<CT> *lds(int off) {
    return reinterpret_cast<C2<CT> *>(spec_smem) + off;
}

struct AxisPlan {
    int L;
    int npass;  // -1: L has a prime factor above 5
    int radix[8];
};

// Radices of a transform length: 3s and 5s first, then 16s, then one of 8/4/2. The
// first pass writes rows R apart; with an odd radix there those writes spread over
// the LDS banks (rows 16 apart all fall on the same banks).
__host__ __device__ constexpr AxisPlan plan_axis(int L) {
    AxisPlan a{L, 0, {0, 0, 0, 0, 0, 0, 0, 0}};
    int n = L;
    while (n % 3 == 0 && a.npass < 8) a.radix[a.npass++] = 3, n /= 3;
    while (n % 5 == 0 && a.npass < 8) a.radix[a.npass++] = 5, n /= 5;
    while (n % 16 == 0 && a.npass < 8) a.radix[a.npass++] = 16, n /= 16;
    if (n % 8 == 0 && a.npass < 8) a.radix[a.npass++] = 8, n /= 8;
    if (n % 4 == 0 && a.npass < 8) a.radix[a.npass++] = 4, n /= 4;
    if (n % 2 == 0 && a.npass < 8) a.radix[a.npass++] = 2, n /= 2;
    if (n != 1) a.npass = -1;
    return a;
}

// Tile width 2^lgb of a kernel kind: the largest with L·2^lgb ≤ kind_tile (for the z
// kernels also three tiles and the twiddles within SPEC_LDS_Z, down to one column).
template <typename CT>
__host__ __device__ constexpr int tile_lgb(int L, int kind) {
    const int tile = kind_tile<CT>(kind);
    int lg = 0;
    while ((L << (lg + 1)) <= tile) lg++;
    if (kind == KIND_Z)
        while (lg > 0 && sizeof(CT) * 2 * (3 * ((size_t)L << lg) + (size_t)L) > SPEC_LDS_Z) lg--;
    return lg;
}

// ---------------------------------------------------------------------------
// Small DFTs in registers, X_k = Σ_n x_n e^{S·2πi nk/R} (S = -1 forward, +1 inverse).

template <int S, typename CT>
__device__ __forceinline__ void dft2(C2<CT> (&v)[2]) {
    const C2<CT> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int S, typename CT>
__device__ __forceinline__ void dft3(C2<CT> (&v)[3]) {
    constexpr double h = 0.86602540378443864676;  // √3/2
    const C2<CT> t1 = cadd(v[1], v[2]);
    const C2<CT> t2 = csub(v[0], cscale(t1, (CT)0.5));
    const C2<CT> t3 = mul_si<S>(cscale(csub(v[1], v[2]), (CT)h));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
}

template <int S, typename CT>
__device__ __forceinline__ void dft4(C2<CT> (&v)[4]) {
    const C2<CT> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    const C2<CT> c = cadd(v[1], v[3]), d = mul_si<S>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
}

template <int S, typename CT>
__device__ __forceinline__ void dft5(C2<CT> (&v)[5]) {
    constexpr double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;  // cos 2π/5, cos 4π/5
    constexpr double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;   // sin 2π/5, sin 4π/5
    const C2<CT> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const C2<CT> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const C2<CT> x0 = v[0];
    const C2<CT> p1 = cadd(x0, cadd(cscale(a1, (CT)c1), cscale(a2, (CT)c2)));
    const C2<CT> p2 = cadd(x0, cadd(cscale(a1, (CT)c2), cscale(a2, (CT)c1)));
    const C2<CT> q1 = mul_si<S>(cadd(cscale(b1, (CT)s1), cscale(b2, (CT)s2)));
    const C2<CT> q2 = mul_si<S>(csub(cscale(b1, (CT)s2), cscale(b2, (CT)s1)));
    v[0] = cadd(x0, cadd(a1, a2));
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
}

template <int S, typename CT>
__device__ __forceinline__ void dft8(C2<CT> (&v)[8]) {
    constexpr double r = 0.70710678118654752440;  // √½
    C2<CT> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft4<S>(e);
    dft4<S>(o);
    // ω8^k = e^{S·iπk/4}
    const C2<CT> t0 = o[0];
    const C2<CT> t1 = cscale(cadd(o[1], mul_si<S>(o[1])), (CT)r);
    const C2<CT> t2 = mul_si<S>(o[2]);
    const C2<CT> t3 = cscale(csub(mul_si<S>(o[3]), o[3]), (CT)r);
    v[0] = cadd(e[0], t0);
    v[4] = csub(e[0], t0);
    v[1] = cadd(e[1], t1);
    v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2);
    v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3);
    v[7] = csub(e[3], t3);
}

template <int S, typename CT>
__device__ __forceinline__ void dft16(C2<CT> (&v)[16]) {
    // cos(πk/8), sin(πk/8), k = 0..7
    constexpr double cs[8] = {1.0,
                              0.92387953251128675613,
                              0.70710678118654752440,
                              0.38268343236508977173,
                              0.0,
                              -0.38268343236508977173,
                              -0.70710678118654752440,
                              -0.92387953251128675613};
    constexpr double sn[8] = {0.0,
                              0.38268343236508977173,
                              0.70710678118654752440,
                              0.92387953251128675613,
                              1.0,
                              0.92387953251128675613,
                              0.70710678118654752440,
                              0.38268343236508977173};
    C2<CT> e[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        e[k] = v[2 * k];
        o[k] = v[2 * k + 1];
    }
    dft8<S>(e);
    dft8<S>(o);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const C2<CT> w{(CT)cs[k], (CT)(S * sn[k])};
        const C2<CT> t = k == 0 ? o[0] : cmul(o[k], w);
        v[k] = cadd(e[k], t);
        v[k + 8] = csub(e[k], t);
    }
}

template <int R, int S, typename CT>
__device__ __forceinline__ void dft(C2<CT> (&v)[R]) {
    if constexpr (R == 2) dft2<S>(v);
    if constexpr (R == 3) dft3<S>(v);
    if constexpr (R == 4) dft4<S>(v);
    if constexpr (R == 5) dft5<S>(v);
    if constexpr (R == 8) dft8<S>(v);
    if constexpr (R == 16) dft16<S>(v);
}

// One Stockham pass of radix R on the LDS tile S (L rows × 2^lgB columns),
// Ns = product of the radices of the previous passes; tw[n] = e^{-2πin/L}.
// Kept out of line: inlined into the radix switch, the passes' live ranges
// merge and the kernels need ~50% more registers.
template <typename CT, int R, int S, int KIND>
__device__ __noinline__ void stockham_pass(int s_off, int lgB, int ld, int L, int Ns, int tw_off) {
    C2<CT> *Sm = lds<CT>(s_off);
    const C2<CT> *tw = lds<CT>(tw_off);
    constexpr int NT = kind_nt(KIND), TILE = kind_tile<CT>(KIND);
    constexpr int ITEMS = (TILE / R + NT - 1) / NT;
    const int nb = L / R;
    const int nwork = nb << lgB;
    const int B1 = (1 << lgB) - 1;
    const int twstride = L / (Ns * R);
    C2<CT> v[ITEMS][R];
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const int w = threadIdx.x + NT * i;
        if (w < nwork) {
            const int b = w & B1, j = w >> lgB;
#pragma unroll
            for (int k = 0; k < R; k++) v[i][k] = Sm[(j + k * nb) * ld + b];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const int w = threadIdx.x + NT * i;
        if (w < nwork) {
            const int b = w & B1, j = w >> lgB;
            const int q = j / Ns, jm = j - q * Ns;
            if (Ns > 1) {
#pragma unroll
                for (int k = 1; k < R; k++) {
                    C2<CT> t = tw[k * jm * twstride];
                    if (S > 0) t.im = -t.im;
                    v[i][k] = cmul(v[i][k], t);
                }
            }
            dft<R, S>(v[i]);
            const int base = q * Ns * R + jm;
#pragma unroll
            for (int k = 0; k < R; k++) Sm[(base + k * Ns) * ld + b] = v[i][k];
        }
    }
    __syncthreads();
}

// Whole transform of every column of the tile (natural order in and out).
// s_off, tw_off: element offsets of the tile and of the twiddle table in spec_smem;
// the tile has 2^lgB columns and row stride ld.
template <typename CT, int S, int KIND>
__device__ void lds_fft(int s_off, int lgB, int ld, const AxisPlan &ax, int tw_off) {
    int Ns = 1;
    for (int p = 0; p < ax.npass; p++) {
        const int R = ax.radix[p];
        switch (R) {
            case 16: stockham_pass<CT, 16, S, KIND>(s_off, lgB, ld, ax.L, Ns, tw_off); break;
            case 8: stockham_pass<CT, 8, S, KIND>(s_off, lgB, ld, ax.L, Ns, tw_off); break;
            case 4: stockham_pass<CT, 4, S, KIND>(s_off, lgB, ld, ax.L, Ns, tw_off); break;
            case 2: stockham_pass<CT, 2, S, KIND>(s_off, lgB, ld, ax.L, Ns, tw_off); break;
            case 3: stockham_pass<CT, 3, S, KIND>(s_off, lgB, ld, ax.L, Ns, tw_off); break;
            default: stockham_pass<CT, 5, S, KIND>(s_off, lgB, ld, ax.L, Ns, tw_off); break;
        }
        Ns *= R;
    }
}

// The same pass with the length, tile width and stride known at compile time (the
// grid sizes of the common configurations): strides fold into immediate offsets
// and the index arithmetic into shifts and constant multiplies.
template <typename CT, int R, int S, int NT, int L, int BW, int NS>
__device__ __forceinline__ void stockham_static(int s_off, int tw_off) {
    constexpr int NB = L / R, NWORK = NB * BW, TWS = L / (NS * R);
    constexpr int ITEMS = (NWORK + NT - 1) / NT;
    C2<CT> *Sm = lds<CT>(s_off);
    const C2<CT> *tw = lds<CT>(tw_off);
    C2<CT> v[ITEMS][R];
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const int w = threadIdx.x + NT * i;
        if (NWORK % NT == 0 || w < NWORK) {
            const int b = w % BW, j = w / BW;
#pragma unroll
            for (int k = 0; k < R; k++) v[i][k] = Sm[(j + k * NB) * BW + b];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const int w = threadIdx.x + NT * i;
        if (NWORK % NT == 0 || w < NWORK) {
            const int b = w % BW, j = w / BW;
            const int q = j / NS, jm = j - q * NS;
            if constexpr (NS > 1) {
                const int t1 = jm * TWS;
#pragma unroll
                for (int k = 1; k < R; k++) {
                    C2<CT> t = tw[k * t1];
                    if (S > 0) t.im = -t.im;
                    v[i][k] = cmul(v[i][k], t);
                }
            }
            dft<R, S>(v[i]);
            const int base = q * NS * R + jm;
#pragma unroll
            for (int k = 0; k < R; k++) Sm[(base + k * NS) * BW + b] = v[i][k];
        }
    }
    __syncthreads();
}

// All passes of a compile-time plan on a tile of BW columns (row stride BW).
template <typename CT, int S, int NT, int L, int BW, int P = 0, int NS = 1>
__device__ __forceinline__ void fft_static(int s_off, int tw_off) {
    constexpr AxisPlan rd = plan_axis(L);
    if constexpr (P < rd.npass) {
        constexpr int R = rd.radix[P];
        stockham_static<CT, R, S, NT, L, BW, NS>(s_off, tw_off);
        fft_static<CT, S, NT, L, BW, P + 1, NS * R>(s_off, tw_off);
    }
}

// Transform of every column of NC side-by-side tiles of 2^lgB columns each (row
// stride NC·2^lgB): compile-time length LS (one pass sequence over all NC·2^lgB
// columns), or LS = 0 for the run-time plan ax (tile by tile).
template <typename CT, int S, int KIND, int LS, int NC = 1>
__device__ __forceinline__ void fft_tile(int s_off, int lgB, const AxisPlan &ax, int tw_off) {
    if constexpr (LS > 0) {
        fft_static<CT, S, kind_nt(KIND), LS, NC * (1 << tile_lgb<CT>(LS, KIND))>(s_off, tw_off);
    } else {
        for (int c = 0; c < NC; c++) lds_fft<CT, S, KIND>(s_off + (c << lgB), lgB, NC << lgB, ax, tw_off);
    }
}

// Blocks are dealt to the 8 XCDs round-robin; remap so that consecutive tiles
// (which share cache lines) run on the same XCD. The grid is padded to a
// multiple of 8 blocks; tiles ≥ n are idle.
__device__ __forceinline__ int64_t xcd_remap(int64_t n) {
    const int64_t per = ((int64_t)gridDim.x + 7) / 8;
    return (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
}

template <typename CT, int NT>
__device__ __forceinline__ void load_twiddles(C2<CT> *twl, const C2<CT> *__restrict__ tw, int L) {
    for (int n = threadIdx.x; n < L; n += NT) twl[n] = tw[n];
}

// Tile element e = threadIdx.x + NT·i (i < EMAX) ↔ (row e >> lgB, column e & (B-1)):
// gathers all of a thread's global loads before its LDS stores, so that they are in
// flight together. get(row, col, ok) returns the value (ok: e inside the tile).
// (ld, coff): the tile sits at column offset coff of an LDS array with row stride ld
// (ld = 0: a tile of its own, row stride 2^lgB).
template <typename CT, int KIND, typename F>
__device__ __forceinline__ void tile_load(C2<CT> *Sm, int n, int lgB, F get, int ld = 0, int coff = 0) {
    constexpr int NT = kind_nt(KIND), EMAX = kind_tile<CT>(KIND) / NT;
    C2<CT> t[EMAX];
#pragma unroll
    for (int i = 0; i < EMAX; i++) {
        const int e = threadIdx.x + NT * i;
        if (e < n) t[i] = get(e >> lgB, e & ((1 << lgB) - 1));
    }
#pragma unroll
    for (int i = 0; i < EMAX; i++) {
        const int e = threadIdx.x + NT * i;
        if (e < n) Sm[ld ? (e >> lgB) * ld + coff + (e & ((1 << lgB) - 1)) : e] = t[i];
    }
}

template <typename CT, int KIND, typename F>
__device__ __forceinline__ void tile_store(const C2<CT> *Sm, int n, int lgB, F put, int ld = 0, int coff = 0) {
    constexpr int NT = kind_nt(KIND), EMAX = kind_tile<CT>(KIND) / NT;
#pragma unroll
    for (int i = 0; i < EMAX; i++) {
        const int e = threadIdx.x + NT * i;
        if (e < n) {
            const int r = e >> lgB, b = e & ((1 << lgB) - 1);
            put(r, b, Sm[ld ? r * ld + coff + b : e]);
        }
    }
}

// Column-pair transfers for tiles whose columns are contiguous in memory (tile width
// B ≥ 2, even mode counts): lane element e ↔ (row e >> (lgB-1), columns 2·(e mod B/2)
// and the next one), 16-byte accesses in FP32. get2(row, b) / put2(row, b, v) move
// the pair (b, b+1).
template <typename CT, int KIND, typename F>
__device__ __forceinline__ void tile_load2(C2<CT> *Sm, int n, int lgB, F get2, int ld = 0, int coff = 0) {
    constexpr int NT = kind_nt(KIND), EMAX = kind_tile<CT>(KIND) / NT / 2;
    const int n2 = n >> 1, lg2 = lgB - 1;
    C2x2<CT> t[EMAX];
#pragma unroll
    for (int i = 0; i < EMAX; i++) {
        const int e = threadIdx.x + NT * i;
        if (e < n2) t[i] = get2(e >> lg2, (e & ((1 << lg2) - 1)) << 1);
    }
#pragma unroll
    for (int i = 0; i < EMAX; i++) {
        const int e = threadIdx.x + NT * i;
        if (e < n2) store_pair(Sm + (ld ? (e >> lg2) * ld + coff + ((e & ((1 << lg2) - 1)) << 1) : 2 * e), t[i]);
    }
}

template <typename CT, int KIND, typename F>
__device__ __forceinline__ void tile_store2(const C2<CT> *Sm, int n, int lgB, F put2, int ld = 0, int coff = 0) {
    constexpr int NT = kind_nt(KIND), EMAX = kind_tile<CT>(KIND) / NT / 2;
    const int n2 = n >> 1, lg2 = lgB - 1;
#pragma unroll
    for (int i = 0; i < EMAX; i++) {
        const int e = threadIdx.x + NT * i;
        if (e < n2) {
            const int r = e >> lg2, b = (e & ((1 << lg2) - 1)) << 1;
            put2(r, b, load_pair(Sm + (ld ? r * ld + coff + b : 2 * e)));
        }
    }
}

// Pair of consecutive complex values at p of which `valid` (≤ 0, 1, ≥ 2) exist.
template <typename CT>
__device__ __forceinline__ C2x2<CT> ld2(const C2<CT> *p, int valid) {
    if (valid >= 2) return load_pair(p);
    C2x2<CT> v{{(CT)0, (CT)0}, {(CT)0, (CT)0}};
    if (valid == 1) v.a = p[0];
    return v;
}
template <typename CT>
__device__ __forceinline__ void st2(C2<CT> *p, const C2x2<CT> &v, int valid) {
    if (valid >= 2)
        store_pair(p, v);
    else if (valid == 1)
        p[0] = v.a;
}

struct SpecGeom {
    int G[3];
    int nkx, nky, nkz, K1, K2;
    int lgB[3];       // log2 of the tile width of the x, y and z kernels
    AxisPlan ax[3];
};

// ---------------------------------------------------------------------------
// Forward x: tile = 2^lgB consecutive pair rows r (r = (c·G2 + z)·G1/2 + p).
// The two real rows y = 2p, 2p + 1 of pair row r are rows 2r, 2r + 1 of x1.
template <typename CT, int LS>
__global__ __launch_bounds__(NTX) void spec_x_forward(SpecGeom g, const C2<CT> *__restrict__ rows,
                                                      C2<CT> *__restrict__ x1, const C2<CT> *__restrict__ tw) {
    const int L = LS ? LS : g.G[0];
    const int lgB = LS ? tile_lgb<CT>(LS, KIND_X) : g.lgB[0];
    const int B = 1 << lgB;
    C2<CT> *Sm = lds<CT>(0);
    const int tw_off = L << lgB;
    C2<CT> *twl = lds<CT>(tw_off);
    const int64_t nrows = 3 * (int64_t)g.G[2] * (g.G[1] >> 1);
    const int64_t ntiles = nrows >> lgB;
    const int64_t tile = xcd_remap(ntiles);
    if (tile >= ntiles) return;
    const int64_t r0 = tile << lgB;
    const C2<CT> *src = rows + r0 * L;
    // S[x·B + b] = row b, column x; two x per lane (rows have even length)
    if constexpr (!use_pairs<CT>()) {
        tile_load<CT, KIND_X>(Sm, L << lgB, lgB, [&](int x, int b) { return src[(int64_t)b * L + x]; });
    } else {
        constexpr int EMAX = kind_tile<CT>(KIND_X) / NTX / 2;
        const int n2 = (L << lgB) >> 1;
        C2x2<CT> t[EMAX];
#pragma unroll
        for (int i = 0; i < EMAX; i++) {
            const int e = threadIdx.x + NTX * i;
            if (e < n2) t[i] = load_pair(src + (int64_t)(e & (B - 1)) * L + (e >> lgB) * 2);
        }
#pragma unroll
        for (int i = 0; i < EMAX; i++) {
            const int e = threadIdx.x + NTX * i;
            if (e < n2) {
                const int b = e & (B - 1), x = (e >> lgB) * 2;
                Sm[(x << lgB) + b] = t[i].a;
                Sm[((x + 1) << lgB) + b] = t[i].b;
            }
        }
    }
    load_twiddles<CT, NTX>(twl, tw, L);
    __syncthreads();
    fft_tile<CT, -1, KIND_X, LS>(0, lgB, g.ax[0], tw_off);
    const int nkx = g.nkx;
    C2<CT> *dst = x1 + 2 * r0 * nkx;
    auto split = [&](int k, int b, C2<CT> &A, C2<CT> &Bv) {
        const C2<CT> zk = Sm[(k << lgB) + b];
        const C2<CT> zm = Sm[((k == 0 ? 0 : L - k) << lgB) + b];
        A = C2<CT>{(zk.re + zm.re) * (CT)0.5, (zk.im - zm.im) * (CT)0.5};
        Bv = C2<CT>{(zk.im + zm.im) * (CT)0.5, (zm.re - zk.re) * (CT)0.5};
    };
    // consecutive lanes take consecutive k of one row: each store instruction writes one
    // contiguous run of x1 (column-fastest lanes halve the LDS bank conflicts of the reads
    // below but split the stores into four runs, and this kernel is bound by HBM)
    if (use_pairs<CT>() && !(nkx & 1)) {
        const int h = nkx >> 1;  // pairs (k, k+1) per row
        for (int e = threadIdx.x; e < (h << lgB); e += NTX) {
            const int b = e / h, k = (e - b * h) * 2;
            C2x2<CT> A, Bv;
            split(k, b, A.a, Bv.a);
            split(k + 1, b, A.b, Bv.b);
            store_pair(dst + (2 * b) * nkx + k, A);
            store_pair(dst + (2 * b + 1) * nkx + k, Bv);
        }
    } else {
        for (int e = threadIdx.x; e < (nkx << lgB); e += NTX) {
            const int b = e / nkx, k = e - b * nkx;
            C2<CT> A, Bv;
            split(k, b, A, Bv);
            dst[(2 * b) * nkx + k] = A;
            dst[(2 * b + 1) * nkx + k] = Bv;
        }
    }
}

// Forward y: tile = (c, z, kx-block); x1[c][z][y][kx] → x2[c][jy][z][kx] for the
// retained ky = jy (jy ≤ K1) or jy - nky.
template <typename CT, int LS>
__global__ __launch_bounds__(NTY) void spec_y_forward(SpecGeom g, const C2<CT> *__restrict__ x1,
                                                      C2<CT> *__restrict__ x2, const C2<CT> *__restrict__ tw) {
    const int L = LS ? LS : g.G[1];
    const int lgB = LS ? tile_lgb<CT>(LS, KIND_Y) : g.lgB[1];
    const int B = 1 << lgB;
    C2<CT> *Sm = lds<CT>(0);
    const int tw_off = L << lgB;
    C2<CT> *twl = lds<CT>(tw_off);
    const int nkx = g.nkx, nkb = (nkx + B - 1) >> lgB;
    const int64_t ntiles = 3 * (int64_t)g.G[2] * nkb;
    const int64_t tile = xcd_remap(ntiles);
    if (tile >= ntiles) return;
    const int kb = (int)(tile % nkb);
    const int64_t cz = tile / nkb;  // c·G2 + z
    const int c = (int)(cz / g.G[2]), z = (int)(cz - (int64_t)c * g.G[2]);
    const int kx0 = kb << lgB;
    const C2<CT> *src = x1 + cz * L * nkx + kx0;
    const bool pairs = use_pairs<CT>() && lgB >= 1 && !(nkx & 1);
    if (pairs)
        tile_load2<CT, KIND_Y>(Sm, L << lgB, lgB,
                               [&](int y, int b) { return ld2(src + (int64_t)y * nkx + b, nkx - kx0 - b); });
    else
        tile_load<CT, KIND_Y>(Sm, L << lgB, lgB, [&](int y, int b) {
            return kx0 + b < nkx ? src[(int64_t)y * nkx + b] : C2<CT>{(CT)0, (CT)0};
        });
    load_twiddles<CT, NTY>(twl, tw, L);
    __syncthreads();
    fft_tile<CT, -1, KIND_Y, LS>(0, lgB, g.ax[1], tw_off);
    const int nky = g.nky, K1 = g.K1;
    C2<CT> *dst = x2 + ((int64_t)c * nky * g.G[2] + z) * nkx + kx0;
    const int64_t jstride = (int64_t)g.G[2] * nkx;
    if (pairs) {
        constexpr int EMAX = kind_tile<CT>(KIND_Y) / NTY / 2;
        const int lg2 = lgB - 1;
#pragma unroll
        for (int i = 0; i < EMAX; i++) {
            const int e = threadIdx.x + NTY * i;
            const int b = (e & ((1 << lg2) - 1)) << 1, jy = e >> lg2;
            if (jy < nky) {
                const int iy = jy <= K1 ? jy : L - nky + jy;
                st2(dst + jy * jstride + b, load_pair(Sm + (iy << lgB) + b), nkx - kx0 - b);
            }
        }
    } else {
        constexpr int EMAX = kind_tile<CT>(KIND_Y) / NTY;
#pragma unroll
        for (int i = 0; i < EMAX; i++) {
            const int e = threadIdx.x + NTY * i;
            const int b = e & (B - 1), jy = e >> lgB;
            if (jy < nky && kx0 + b < nkx) {
                const int iy = jy <= K1 ? jy : L - nky + jy;
                dst[jy * jstride + b] = Sm[(iy << lgB) + b];
            }
        }
    }
}

struct SpecFourier {
    double kscale[3];
    double pref;          // Γ/V
    double kmax2_trunc;   // < 0: no spherical truncation
    const double *dx, *dy, *dz;  // 1/ψ̂ (deconvolution), dy/dz indexed by k + K
    const double *ex, *ey, *ez;  // (1/ψ̂)² e^{-k_d²/4α²}, same indexing
    int do_vel, do_psi, keep_uhat;
    void *uhat;                  // [3][nky][nkz][nkx]
    int64_t ucstride;            // complex values per component
};

// z passes fused with the Fourier-space kernel: tile = (jy, kx-block), the three
// components in three LDS tiles. x2 holds z rows on input; on output, x2 the
// velocity and x2b the streamfunction (each when requested), in z space, ready
// for the inverse y, x passes. With both outputs, the velocity spectrum makes a
// round trip through x2 while the tiles transform the streamfunction.
template <typename CT, int LS>
__global__ __launch_bounds__(NTZ) void spec_z_fourier(SpecGeom g, SpecFourier f, C2<CT> *__restrict__ x2,
                                                      C2<CT> *__restrict__ x2b, const C2<CT> *__restrict__ tw) {
    constexpr int EMAX = kind_tile<CT>(KIND_Z) / NTZ;
    const int L = LS ? LS : g.G[2];
    const int lgB = LS ? tile_lgb<CT>(LS, KIND_Z) : g.lgB[2];
    const int B = 1 << lgB;
    const int n = L << lgB;
    // the three component tiles side by side: component c at columns [c·B, c·B + B) of
    // an LDS array with row stride 3B, so that one transform call covers all three
    C2<CT> *Sm = lds<CT>(0);
    const int ld = 3 << lgB;
    const int tw_off = 3 * n;
    C2<CT> *twl = lds<CT>(tw_off);
    const int nkx = g.nkx, nky = g.nky, nkb = (nkx + B - 1) >> lgB;
    const int64_t ntiles = (int64_t)nky * nkb;
    const int64_t tile = xcd_remap(ntiles);
    if (tile >= ntiles) return;
    const int kb = (int)(tile % nkb), jy = (int)(tile / nkb);
    const int kx0 = kb << lgB;
    const int64_t cs = (int64_t)nky * L * nkx;  // complex values per component of x2
    const int64_t off = (int64_t)jy * L * nkx + kx0;
    load_twiddles<CT, NTZ>(twl, tw, L);
    const bool pairs = use_pairs<CT>() && lgB >= 1 && !(nkx & 1);
    auto load3 = [&]() {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const C2<CT> *src = x2 + c * cs + off;
            if (pairs)
                tile_load2<CT, KIND_Z>(
                    Sm, n, lgB, [&](int z, int b) { return ld2(src + (int64_t)z * nkx + b, nkx - kx0 - b); }, ld,
                    c << lgB);
            else
                tile_load<CT, KIND_Z>(
                    Sm, n, lgB,
                    [&](int z, int b) { return kx0 + b < nkx ? src[(int64_t)z * nkx + b] : C2<CT>{(CT)0, (CT)0}; },
                    ld, c << lgB);
        }
    };
    auto store3 = [&](C2<CT> *out) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            C2<CT> *dst = out + c * cs + off;
            if (pairs)
                tile_store2<CT, KIND_Z>(
                    Sm, n, lgB,
                    [&](int z, int b, const C2x2<CT> &v) { st2(dst + (int64_t)z * nkx + b, v, nkx - kx0 - b); }, ld,
                    c << lgB);
            else
                tile_store<CT, KIND_Z>(
                    Sm, n, lgB,
                    [&](int z, int b, C2<CT> v) {
                        if (kx0 + b < nkx) dst[(int64_t)z * nkx + b] = v;
                    },
                    ld, c << lgB);
        }
    };
    load3();
    __syncthreads();
    fft_tile<CT, -1, KIND_Z, LS, 3>(0, lgB, g.ax[2], tw_off);
    // Fourier-space operators on the retained modes
    const int K1 = g.K1, K2 = g.K2, nkz = g.nkz;
    const int ky = jy <= K1 ? jy : jy - nky;
    const double KY = f.kscale[1] * ky;
    const bool both = f.do_vel && f.do_psi;
    // a thread's elements share one column (NTZ is a multiple of B): the kx and ky
    // factors are loaded once, the kz factors all at once before the loop
    const double fxy = threadIdx.x < n && kx0 + (threadIdx.x & (B - 1)) < nkx
                           ? f.pref * f.ex[kx0 + (threadIdx.x & (B - 1))] * f.ey[ky + K1]
                           : 0.0;
    double fz[EMAX];
#pragma unroll
    for (int i = 0; i < EMAX; i++) {
        const int e = threadIdx.x + NTZ * i;
        const int iz = e >> lgB, kz = iz <= (L >> 1) ? iz : iz - L;
        fz[i] = (e < n && kz >= -K2 && kz <= K2) ? f.ez[kz + K2] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < EMAX; i++) {
        const int e = threadIdx.x + NTZ * i;
        if (e >= n) break;
        const int b = e & (B - 1), iz = e >> lgB;
        C2<CT> *T0 = Sm + iz * ld + b;  // component d at T0[d << lgB]
        const int kx = kx0 + b;
        const int kz = iz <= (L >> 1) ? iz : iz - L;
        const bool kz_kept = kz >= -K2 && kz <= K2;
        const double KX = f.kscale[0] * kx, KZ = f.kscale[2] * kz;
        const double k2 = KX * KX + KY * KY + KZ * KZ;
        bool keep = kx < nkx && kz_kept && k2 != 0;
        if (f.kmax2_trunc >= 0 && k2 > f.kmax2_trunc) keep = false;
        C2<CT> *u = nullptr;
        if (f.keep_uhat && kx < nkx && kz_kept) {
            const int jz = kz >= 0 ? kz : kz + nkz;
            u = (C2<CT> *)f.uhat + ((int64_t)jy * nkz + jz) * nkx + kx;
        }
        const C2<CT> zero{(CT)0, (CT)0};
        if (!keep) {
            if (u) {
#pragma unroll
                for (int d = 0; d < 3; d++) u[d * f.ucstride] = zero;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) T0[d << lgB] = zero;
            if (both && kx < nkx) {
#pragma unroll
                for (int d = 0; d < 3; d++) x2[d * cs + off + (int64_t)iz * nkx + b] = zero;
            }
            continue;
        }
        double wr[3], wi[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const C2<CT> w = T0[d << lgB];
            wr[d] = (double)w.re;
            wi[d] = (double)w.im;
        }
        const double inv_k2 = 1.0 / k2;
        if (u) {
            // unsmoothed field ω̂/k² (one deconvolution, no Ewald factor), then i k × for the velocity
            const double su = f.pref * f.dx[kx] * f.dy[ky + K1] * f.dz[kz + K2] * inv_k2;
            double ur[3], ui[3];
#pragma unroll
            for (int d = 0; d < 3; d++) ur[d] = wr[d] * su, ui[d] = wi[d] * su;
            if (f.do_vel) {
                const double ax = KY * ur[2] - KZ * ur[1], bx = KY * ui[2] - KZ * ui[1];
                const double ay = KZ * ur[0] - KX * ur[2], by = KZ * ui[0] - KX * ui[2];
                const double az = KX * ur[1] - KY * ur[0], bz = KX * ui[1] - KY * ui[0];
                ur[0] = -bx, ui[0] = ax, ur[1] = -by, ui[1] = ay, ur[2] = -bz, ui[2] = az;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) u[d * f.ucstride] = C2<CT>{(CT)ur[d], (CT)ui[d]};
        }
        // ω̂ = Γ/V · B̂/ψ̂; ψ̂_out = ω̂·e^{-k²/4α²}/(ψ̂ k²) (deconvolution of type 1 and type 2)
        const double s = fxy * fz[i] * inv_k2;
#pragma unroll
        for (int d = 0; d < 3; d++) wr[d] *= s, wi[d] *= s;
        if (f.do_psi) {
#pragma unroll
            for (int d = 0; d < 3; d++) T0[d << lgB] = C2<CT>{(CT)wr[d], (CT)wi[d]};
        }
        if (f.do_vel) {
            // v̂ = i k × ψ̂: (i k × w)_x = i (KY w_z - KZ w_y), i (a + ib) = -b + ia
            const double ax = KY * wr[2] - KZ * wr[1], bx = KY * wi[2] - KZ * wi[1];
            const double ay = KZ * wr[0] - KX * wr[2], by = KZ * wi[0] - KX * wi[2];
            const double az = KX * wr[1] - KY * wr[0], bz = KX * wi[1] - KY * wi[0];
            const C2<CT> v[3] = {{(CT)(-bx), (CT)ax}, {(CT)(-by), (CT)ay}, {(CT)(-bz), (CT)az}};
            if (both) {
#pragma unroll
                for (int d = 0; d < 3; d++) x2[d * cs + off + (int64_t)iz * nkx + b] = v[d];
            } else {
#pragma unroll
                for (int d = 0; d < 3; d++) T0[d << lgB] = v[d];
            }
        }
    }
    __syncthreads();
    fft_tile<CT, 1, KIND_Z, LS, 3>(0, lgB, g.ax[2], tw_off);
    store3(f.do_psi ? x2b : x2);
    if (both) {
        // this block's writes to x2 above are visible to its own loads after the barrier
        __threadfence_block();
        __syncthreads();
        load3();
        __syncthreads();
        fft_tile<CT, 1, KIND_Z, LS, 3>(0, lgB, g.ax[2], tw_off);
        store3(x2);
    }
}

// Band-filtered kept field in z space: the kept û on kmin² ≤ k² ≤ kmax² times the
// type-2 deconvolution, zero elsewhere, inverse FFT_z → x2. Tile = (jy, kx-block).
// With cf > 0 the band is further weighted by e^{−cf·k²} − e^{−cc·k²} (difference of two
// Ewald Gaussians, cf = 1/4α₁², cc = 1/4α₂²).
template <typename CT, int LS>
__global__ __launch_bounds__(NTZ) void spec_z_band(SpecGeom g, SpecFourier f, double kmin2, double kmax2, double cf,
                                                   double cc, C2<CT> *__restrict__ x2, const C2<CT> *__restrict__ tw) {
    const int L = LS ? LS : g.G[2];
    const int lgB = LS ? tile_lgb<CT>(LS, KIND_Z) : g.lgB[2];
    const int B = 1 << lgB;
    C2<CT> *Sm = lds<CT>(0);
    const int tw_off = L << lgB;
    C2<CT> *twl = lds<CT>(tw_off);
    const int nkx = g.nkx, nky = g.nky, nkz = g.nkz, nkb = (nkx + B - 1) >> lgB;
    const int64_t ntiles = (int64_t)nky * nkb;
    const int64_t tile = xcd_remap(ntiles);
    if (tile >= ntiles) return;
    const int kb = (int)(tile % nkb), jy = (int)(tile / nkb);
    const int kx0 = kb << lgB;
    const int n = L << lgB;
    const int64_t cs = (int64_t)nky * L * nkx;
    const int64_t off = (int64_t)jy * L * nkx + kx0;
    const int K1 = g.K1, K2 = g.K2;
    const int ky = jy <= K1 ? jy : jy - nky;
    const double KY = f.kscale[1] * ky;
    load_twiddles<CT, NTZ>(twl, tw, L);
    const C2<CT> *uh = (const C2<CT> *)f.uhat;
    for (int c = 0; c < 3; c++) {
        tile_load<CT, KIND_Z>(Sm, n, lgB, [&](int iz, int b) {
            const int kx = kx0 + b;
            const int kz = iz <= (L >> 1) ? iz : iz - L;
            C2<CT> v{(CT)0, (CT)0};
            if (kx < nkx && kz >= -K2 && kz <= K2) {
                const double KX = f.kscale[0] * kx, KZ = f.kscale[2] * kz;
                const double k2 = KX * KX + KY * KY + KZ * KZ;
                if (k2 >= kmin2 && k2 <= kmax2) {
                    double D = f.dx[kx] * f.dy[ky + K1] * f.dz[kz + K2];
                    if (cf > 0) D *= exp(-cf * k2) - exp(-cc * k2);
                    const int jz = kz >= 0 ? kz : kz + nkz;
                    const C2<CT> u = uh[c * f.ucstride + ((int64_t)jy * nkz + jz) * nkx + kx];
                    v = C2<CT>{(CT)((double)u.re * D), (CT)((double)u.im * D)};
                }
            }
            return v;
        });
        __syncthreads();
        fft_tile<CT, 1, KIND_Z, LS>(0, lgB, g.ax[2], tw_off);
        C2<CT> *dst = x2 + c * cs + off;
        tile_store<CT, KIND_Z>(Sm, n, lgB, [&](int z, int b, C2<CT> v) {
            if (kx0 + b < nkx) dst[(int64_t)z * nkx + b] = v;
        });
        __syncthreads();
    }
}

// Inverse y: tile = (c, z, kx-block); x2[c][jy][z][kx] on the retained ky (zero
// elsewhere) → inverse FFT_y → x1[c][z][y][kx].
template <typename CT, int LS>
__global__ __launch_bounds__(NTY) void spec_y_inverse(SpecGeom g, const C2<CT> *__restrict__ x2,
                                                      C2<CT> *__restrict__ x1, const C2<CT> *__restrict__ tw) {
    const int L = LS ? LS : g.G[1];
    const int lgB = LS ? tile_lgb<CT>(LS, KIND_Y) : g.lgB[1];
    const int B = 1 << lgB;
    C2<CT> *Sm = lds<CT>(0);
    const int tw_off = L << lgB;
    C2<CT> *twl = lds<CT>(tw_off);
    const int nkx = g.nkx, nky = g.nky, K1 = g.K1, nkb = (nkx + B - 1) >> lgB;
    const int64_t ntiles = 3 * (int64_t)g.G[2] * nkb;
    const int64_t tile = xcd_remap(ntiles);
    if (tile >= ntiles) return;
    const int kb = (int)(tile % nkb);
    const int64_t cz = tile / nkb;
    const int c = (int)(cz / g.G[2]), z = (int)(cz - (int64_t)c * g.G[2]);
    const int kx0 = kb << lgB;
    const C2<CT> *src = x2 + ((int64_t)c * nky * g.G[2] + z) * nkx + kx0;
    const int64_t jstride = (int64_t)g.G[2] * nkx;
    const bool pairs = use_pairs<CT>() && lgB >= 1 && !(nkx & 1);
    if (pairs)
        tile_load2<CT, KIND_Y>(Sm, L << lgB, lgB, [&](int iy, int b) {
            const int ky = iy <= (L >> 1) ? iy : iy - L;
            const bool kept = ky >= -K1 && ky <= K1;
            return ld2(src + (kept ? (ky >= 0 ? ky : ky + nky) * jstride + b : 0), kept ? nkx - kx0 - b : 0);
        });
    else
        tile_load<CT, KIND_Y>(Sm, L << lgB, lgB, [&](int iy, int b) {
            const int ky = iy <= (L >> 1) ? iy : iy - L;
            C2<CT> v{(CT)0, (CT)0};
            if (ky >= -K1 && ky <= K1 && kx0 + b < nkx) v = src[(ky >= 0 ? ky : ky + nky) * jstride + b];
            return v;
        });
    load_twiddles<CT, NTY>(twl, tw, L);
    __syncthreads();
    fft_tile<CT, 1, KIND_Y, LS>(0, lgB, g.ax[1], tw_off);
    C2<CT> *dst = x1 + cz * L * nkx + kx0;
    if (pairs)
        tile_store2<CT, KIND_Y>(Sm, L << lgB, lgB, [&](int y, int b, const C2x2<CT> &v) {
            st2(dst + (int64_t)y * nkx + b, v, nkx - kx0 - b);
        });
    else
        tile_store<CT, KIND_Y>(Sm, L << lgB, lgB, [&](int y, int b, C2<CT> v) {
            if (kx0 + b < nkx) dst[(int64_t)y * nkx + b] = v;
        });
}

// Inverse x: tile = 2^lgB pair rows; Z_k = A_k + i B_k for k < nkx (A, B from rows
// 2r, 2r + 1 of x1), their conjugates at G0 - k (Hermitian symmetry of real rows),
// zero otherwise, DC keeps only its real parts; inverse FFT_x → pair rows.
template <typename CT, int LS>
__global__ __launch_bounds__(NTX) void spec_x_inverse(SpecGeom g, const C2<CT> *__restrict__ x1,
                                                      C2<CT> *__restrict__ rows, const C2<CT> *__restrict__ tw) {
    const int L = LS ? LS : g.G[0];
    const int lgB = LS ? tile_lgb<CT>(LS, KIND_X) : g.lgB[0];
    const int B = 1 << lgB;
    C2<CT> *Sm = lds<CT>(0);
    const int tw_off = L << lgB;
    C2<CT> *twl = lds<CT>(tw_off);
    const int64_t nrows = 3 * (int64_t)g.G[2] * (g.G[1] >> 1);
    const int64_t ntiles = nrows >> lgB;
    const int64_t tile = xcd_remap(ntiles);
    if (tile >= ntiles) return;
    const int64_t r0 = tile << lgB;
    const int nkx = g.nkx;
    const C2<CT> *src = x1 + 2 * r0 * nkx;
    // each (A_k, B_k) is read once and gives both Z_k and Z_{G0-k}
    {
        auto put = [&](int k, int b, C2<CT> A, C2<CT> Bv) {
            if (k == 0) A.im = Bv.im = (CT)0;
            Sm[(k << lgB) + b] = C2<CT>{A.re - Bv.im, A.im + Bv.re};
            if (k > 0) Sm[((L - k) << lgB) + b] = C2<CT>{A.re + Bv.im, Bv.re - A.im};  // conj(A) + i conj(B)
        };
        if (use_pairs<CT>() && !(nkx & 1)) {
            // pairs (k, k+1) per row: 16-byte loads in FP32
            constexpr int EMAX = kind_tile<CT>(KIND_X) / NTX / 2;
            const int nin = (nkx >> 1) << lgB;
            C2x2<CT> a[EMAX], bb[EMAX];
#pragma unroll
            for (int i = 0; i < EMAX; i++) {
                const int e = threadIdx.x + NTX * i;
                if (e < nin) {
                    const int k = (e >> lgB) * 2, b = e & (B - 1);
                    a[i] = load_pair(src + (2 * b) * nkx + k);
                    bb[i] = load_pair(src + (2 * b + 1) * nkx + k);
                }
            }
#pragma unroll
            for (int i = 0; i < EMAX; i++) {
                const int e = threadIdx.x + NTX * i;
                if (e < nin) {
                    const int k = (e >> lgB) * 2, b = e & (B - 1);
                    put(k, b, a[i].a, bb[i].a);
                    put(k + 1, b, a[i].b, bb[i].b);
                }
            }
        } else {
            constexpr int EMAX = kind_tile<CT>(KIND_X) / NTX;
            const int nin = nkx << lgB;
            C2<CT> a[EMAX], bb[EMAX];
#pragma unroll
            for (int i = 0; i < EMAX; i++) {
                const int e = threadIdx.x + NTX * i;
                if (e < nin) {
                    const int k = e >> lgB, b = e & (B - 1);
                    a[i] = src[(2 * b) * nkx + k];
                    bb[i] = src[(2 * b + 1) * nkx + k];
                }
            }
#pragma unroll
            for (int i = 0; i < EMAX; i++) {
                const int e = threadIdx.x + NTX * i;
                if (e < nin) put(e >> lgB, e & (B - 1), a[i], bb[i]);
            }
        }
        // modes kx ∈ [nkx, G0 - nkx] are not retained
        for (int e = (nkx << lgB) + threadIdx.x; e < ((L - nkx + 1) << lgB); e += NTX) Sm[e] = C2<CT>{(CT)0, (CT)0};
    }
    load_twiddles<CT, NTX>(twl, tw, L);
    __syncthreads();
    fft_tile<CT, 1, KIND_X, LS>(0, lgB, g.ax[0], tw_off);
    // rows out two values per lane (16-byte stores in FP32), lanes along b then x
    C2<CT> *dst = rows + r0 * L;
    if constexpr (!use_pairs<CT>()) {
        tile_store<CT, KIND_X>(Sm, L << lgB, lgB, [&](int x, int b, C2<CT> v) { dst[(int64_t)b * L + x] = v; });
    } else {
        constexpr int NT = NTX, EMAX = kind_tile<CT>(KIND_X) / NT / 2;
        const int n2 = (L << lgB) >> 1;
#pragma unroll
        for (int i = 0; i < EMAX; i++) {
            const int e = threadIdx.x + NT * i;
            if (e < n2) {
                const int b = e & (B - 1), x = (e >> lgB) * 2;
                C2x2<CT> v;
                v.a = Sm[(x << lgB) + b];
                v.b = Sm[((x + 1) << lgB) + b];
                store_pair(dst + (int64_t)b * L + x, v);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Inverse x transform fused with the interpolation (type 2).
//
// Slab = SLAB_HY y-rows × SLAB_HZ z-rows × all G0 x-cells of the 3 components,
// held in LDS as SLAB_BW complex columns b = (c·SLAB_HZ + zz)·SLAB_HP + pp (pair
// row pp of z-row zz, component c) at S[x·SLAB_BW + b], then the twiddles. The block rebuilds the
// columns from x1 as spec_x_inverse does and transforms them: S then holds the
// grid values of the slab rows (real part: even y, imaginary part: odd y). Every
// node whose W³ stencil reaches the slab has its first stencil cell in one of
// NYH × NZH home slabs (nodes come sorted by home slab); eight lanes per node,
// one per slab row, sum their row's W cells with the x-weights, scale by the
// row's y- and z-weights, and a butterfly over the eight gives the part of the
// node's stencil sum that lies in this slab. It goes to slot q of the node, q =
// the home-slab range it was found in (one slab per slot); slab_reduce adds the
// slots in a fixed order. No grid written, no atomics: deterministic.
constexpr int SLAB_HP = SLAB_HY / 2;
constexpr int SLAB_ROWS = SLAB_HY * SLAB_HZ;
constexpr int SLAB_NT = 768;  // 768-cell slabs: 4 radix-3 butterflies per thread, then at most 1 radix-16
constexpr int SLAB_BW = 3 * SLAB_HZ * SLAB_HP;
static_assert(SLAB_ROWS == 8 && SLAB_HY % 2 == 0, "eight slab rows of whole pair rows: eight lanes per node");

// Home slabs along y (z) of the nodes whose W-cell stencil reaches a slab.
constexpr int slab_nyh(int W) { return 1 + (W - 1 + SLAB_HY - 1) / SLAB_HY; }
constexpr int slab_nzh(int W) { return 1 + (W - 1 + SLAB_HZ - 1) / SLAB_HZ; }
// Node record: W x-weights, W y-weights, W z-weights, padded to whole 16-byte vectors.
template <typename CT>
constexpr int slab_rec(int W) {
    return slab_rec_len(W, (int)sizeof(CT));
}

// Lane-to-lane move by a DPP control (every lane has a source), FP32 and FP64.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over each group of 8 consecutive lanes, in every lane of the group: partners at xor 1
// and xor 2 (quad permutations), then the mirror lane of the half-row, which after those
// steps holds the same value as the xor-4 partner: the additions of an xor butterfly over
// offsets 1, 2, 4, so the same bits, without going through LDS.
template <typename CT>
__device__ __forceinline__ CT sum8_lanes(CT v) {
    v += dpp_move<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_move<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_move<0x141>(v);  // row_half_mirror
    return v;
}

// i in (-n, 2n) -> i mod n
__device__ __forceinline__ int wrap1(int i, int n) {
    i += (i < 0) ? n : 0;
    i -= (i >= n) ? n : 0;
    return i;
}

struct SlabArgs {
    int G[3];
    int nkx;
    int64_t Np;
    const int32_t *start;  // start[sz·nsy + sy]: first node of home slab (sy, sz); start[nsy·nsz] = Np
    const int32_t *si0;    // [3][Np] first stencil cell per direction of the slab-sorted nodes, in [0, G_d)
    const void *rec;       // [Np][slab_rec(W)] node records (compute type)
    void *part;            // [NYH·NZH][3][Np] stencil sums per slot (compute type)
};

template <typename CT, int W, int LS>
__global__ __launch_bounds__(SLAB_NT) void x_inverse_interp(SlabArgs a, const C2<CT> *__restrict__ x1,
                                                            const C2<CT> *__restrict__ tw) {
    constexpr int REC = slab_rec<CT>(W);
    constexpr int NYH = slab_nyh(W), NZH = slab_nzh(W), NRG = NYH * NZH;
    static_assert(NRG <= 64, "one lane per home slab");
    constexpr int CST = 2 * SLAB_HZ * SLAB_HP;  // values between the columns of consecutive components
    const int G1 = a.G[1], G2 = a.G[2];
    const int nsy = G1 / SLAB_HY, nsz = G2 / SLAB_HZ;
    const int64_t nslab = (int64_t)nsy * nsz;
    const int64_t sl = xcd_remap(nslab);
    if (sl >= nslab) return;
    // slabs go in 8×8 blocks (ty fastest): the slabs in flight on one XCD share x1 rows and nodes
    const int blk = (int)(sl >> 6), in = (int)(sl & 63), nby = nsy >> 3;
    const int ty = ((blk % nby) << 3) + (in & 7), tz = ((blk / nby) << 3) + (in >> 3);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    // nodes: the home-slab ranges, concatenated (lane r < NRG: range r)
    int s0 = 0, len = 0;
    if (lane < NRG) {
        const int sy = wrap1(ty - (NYH - 1) + lane % NYH, nsy);
        const int sz = wrap1(tz - (NZH - 1) + lane / NYH, nsz);
        const int key = sz * nsy + sy;
        s0 = a.start[key];
        len = a.start[key + 1] - s0;
    }
    const int incl = wave_incl_scan(len);
    const int ncand = __builtin_amdgcn_readlane(incl, NRG - 1);  // the same in every wavefront
    if (ncand == 0) return;
    // candidate cnd of range q is node off[q] + cnd (cnd ≥ pre[q]); wavefront-uniform
    const int excl = incl - len;
    const int offv = s0 - excl;  // lane q: node of candidate 0 of range q
    int pre[NRG];
#pragma unroll
    for (int q = 0; q < NRG; q++) pre[q] = __builtin_amdgcn_readlane(excl, q);

    // columns: Z_k = A_k + i B_k (k < nkx; DC real) and conj(A_k) + i conj(B_k) at L - k
    C2<CT> *Sm = lds<CT>(0);
    const int nkx = a.nkx;
    auto row_of = [&](int b) {  // x1 row of the even y of column b
        const int c = b / (SLAB_HZ * SLAB_HP), rr = b - c * (SLAB_HZ * SLAB_HP);
        const int zz = rr / SLAB_HP, pp = rr - zz * SLAB_HP;
        return ((int64_t)c * G2 + tz * SLAB_HZ + zz) * G1 + ty * SLAB_HY + 2 * pp;
    };
    auto put = [&](int k, int b, C2<CT> A, C2<CT> Bv) {
        if (k == 0) A.im = Bv.im = (CT)0;
        Sm[k * SLAB_BW + b] = C2<CT>{A.re - Bv.im, A.im + Bv.re};
        if (k > 0) Sm[(LS - k) * SLAB_BW + b] = C2<CT>{A.re + Bv.im, Bv.re - A.im};
    };
    // consecutive lanes take consecutive columns b (LDS writes along a slab row)
    if (use_pairs<CT>() && !(nkx & 1)) {
        const int h = nkx >> 1;  // pairs (k, k+1) per row: 16-byte loads
        for (int e = threadIdx.x; e < h * SLAB_BW; e += SLAB_NT) {
            const int kp = e / SLAB_BW, b = e - kp * SLAB_BW, k = kp * 2;
            const C2<CT> *src = x1 + row_of(b) * nkx + k;
            const C2x2<CT> A = load_pair(src), Bv = load_pair(src + nkx);
            put(k, b, A.a, Bv.a);
            put(k + 1, b, A.b, Bv.b);
        }
    } else {
        for (int e = threadIdx.x; e < nkx * SLAB_BW; e += SLAB_NT) {
            const int k = e / SLAB_BW, b = e - k * SLAB_BW;
            const C2<CT> *src = x1 + row_of(b) * nkx + k;
            put(k, b, src[0], src[nkx]);
        }
    }
    load_twiddles<CT, SLAB_NT>(Sm + LS * SLAB_BW, tw, LS);
    // modes kx ∈ [nkx, L - nkx] are not retained
    for (int e = threadIdx.x; e < (LS - 2 * nkx + 1) * SLAB_BW; e += SLAB_NT) {
        const int k = nkx + e / SLAB_BW, b = e - (e / SLAB_BW) * SLAB_BW;
        Sm[k * SLAB_BW + b] = C2<CT>{(CT)0, (CT)0};
    }
    __syncthreads();
    fft_static<CT, 1, SLAB_NT, LS, SLAB_BW>(0, LS * SLAB_BW);
    // x rows 0..W-2 again after row LS-1 (over the twiddles, no longer needed), so that a
    // node's W x-cells are consecutive rows also across the periodic end
    static_assert((W - 1) * SLAB_BW <= LS, "the x halo fits in the twiddle area");
    for (int e = threadIdx.x; e < (W - 1) * SLAB_BW; e += SLAB_NT) Sm[LS * SLAB_BW + e] = Sm[e];
    __syncthreads();

    // stencil sums: lane (j, r) = node j of the group, slab row r
    const CT *Sf = reinterpret_cast<const CT *>(spec_smem);
    const int j = lane >> 3, r = lane & 7;
    const int yy = r % SLAB_HY, zz = r / SLAB_HY;
    const int y = ty * SLAB_HY + yy, z = tz * SLAB_HZ + zz;
    const CT *row = Sf + (zz * SLAB_HP + (yy >> 1)) * 2 + (yy & 1);  // + 2·SLAB_BW·x + CST·c
    const int64_t Np = a.Np;
    const int32_t *si0 = a.si0;
    const CT *rec = static_cast<const CT *>(a.rec);
    CT *part = static_cast<CT *>(a.part);
    for (int g = wave; g * 8 < ncand; g += SLAB_NT / 64) {
        const int cnd = g * 8 + j;
        // range of the candidate: the number of later ranges starting at or before it (the
        // starts never decrease); its offset from lane q, with every lane active
        int q = 0;
#pragma unroll
        for (int t = 1; t < NRG; t++) q += cnd >= pre[t] ? 1 : 0;
        const int o = __shfl(offv, q, 64);
        int sidx = 0;
        CT v0 = (CT)0, v1 = (CT)0, v2 = (CT)0;
        if (cnd < ncand) {
            sidx = cnd + o;
            const int ix = si0[sidx], iy = si0[Np + sidx], iz = si0[2 * Np + sidx];
            // the record's x weights (whole vectors), and its y and z weights of this row
            constexpr int VL = 16 / sizeof(CT), NXV = (W + VL - 1) / VL;
            static_assert(NXV * VL <= REC, "x-weight vectors inside the record");
            CT wx[NXV * VL];
            const CT *rp = rec + (int64_t)sidx * REC;
            if constexpr (sizeof(CT) == 4) {
#pragma unroll
                for (int k = 0; k < NXV; k++) {
                    const float4 v = reinterpret_cast<const float4 *>(rp)[k];
                    wx[4 * k] = v.x, wx[4 * k + 1] = v.y, wx[4 * k + 2] = v.z, wx[4 * k + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int k = 0; k < NXV; k++) {
                    const double2 v = reinterpret_cast<const double2 *>(rp)[k];
                    wx[2 * k] = v.x, wx[2 * k + 1] = v.y;
                }
            }
            const int ky = wrap1(y - iy, G1), kz = wrap1(z - iz, G2);
            if (ky < W && kz < W) {
                const CT wy = rp[W + ky], wz = rp[2 * W + kz];
                const CT *p = row + ix * (2 * SLAB_BW);  // rows ix .. ix+W-1, the halo past LS-1
#pragma unroll
                for (int k = 0; k < W; k++) {
                    v0 += wx[k] * p[k * 2 * SLAB_BW];
                    v1 += wx[k] * p[k * 2 * SLAB_BW + CST];
                    v2 += wx[k] * p[k * 2 * SLAB_BW + 2 * CST];
                }
                const CT wyz = wy * wz;
                v0 *= wyz, v1 *= wyz, v2 *= wyz;
            }
        }
        v0 = sum8_lanes(v0);
        v1 = sum8_lanes(v1);
        v2 = sum8_lanes(v2);
        if (cnd < ncand && r == 0) {
            CT *dst = part + (int64_t)q * 3 * Np + sidx;
            dst[0] = v0;
            dst[Np] = v1;
            dst[2 * Np] = v2;
        }
    }
}

// out[c][perm[s]] = Σ_q part[q][c][s] over the nslot slots, in slot order.
template <typename CT>
__global__ void slab_reduce(int64_t Np, int nslot, const CT *__restrict__ part, const int32_t *__restrict__ perm,
                            double *__restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Np) return;
    const int64_t i = perm[s];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double acc = 0;
        for (int q = 0; q < nslot; q++) acc += (double)part[((int64_t)q * 3 + c) * Np + s];
        out[c * Np + i] = acc;
    }
}

// ---------------------------------------------------------------------------
// Host side

template <typename CT>
SpecGeom spec_geom(const vpbs_ctx *c) {
    SpecGeom g{};
    for (int d = 0; d < 3; d++) {
        g.G[d] = c->G[d];
        g.ax[d] = plan_axis(c->G[d]);
        g.lgB[d] = tile_lgb<CT>(c->G[d], d);  // kind = axis
    }
    g.nkx = c->nkx;
    g.nky = c->nky;
    g.nkz = c->nkz;
    g.K1 = c->kmax_mode[1];
    g.K2 = c->kmax_mode[2];
    return g;
}

template <typename CT>
size_t spec_lds(const SpecGeom &g, int d, int ntiles = 1) {
    return sizeof(C2<CT>) * (ntiles * ((size_t)g.G[d] << g.lgB[d]) + g.G[d]);
}

unsigned grid8(int64_t ntiles) { return (unsigned)(8 * ((ntiles + 7) / 8)); }

template <typename CT>
const C2<CT> *twiddles(const vpbs_ctx *c, int d) {
    const C2<CT> *t = (const C2<CT> *)c->spec_tw;
    for (int k = 0; k < d; k++) t += c->G[k];
    return t;
}

template <typename CT>
SpecFourier spec_fourier_args(const vpbs_ctx *c) {
    SpecFourier f{};
    for (int d = 0; d < 3; d++) f.kscale[d] = 2 * M_PI / c->p.Ls[d];
    f.pref = c->p.Gamma / (c->p.Ls[0] * c->p.Ls[1] * c->p.Ls[2]);
    f.kmax2_trunc = -1;
    if (c->p.truncate_spherical) {
        double km = 1e300;
        for (int d = 0; d < 3; d++) km = std::fmin(km, 2 * M_PI * ((c->p.Ns[d] - 1) / 2) / c->p.Ls[d]);
        f.kmax2_trunc = km * km;  // params.jl:161-166
    }
    const int n0 = c->kmax_mode[0] + 1, n1 = 2 * c->kmax_mode[1] + 1;
    f.dx = c->deconv;
    f.dy = f.dx + n0;
    f.dz = f.dy + n1;
    f.ex = c->spec_fac;
    f.ey = f.ex + n0;
    f.ez = f.ey + n1;
    f.uhat = c->uhat;
    f.ucstride = c->uhat_stride;
    return f;
}

// Transform lengths compiled with static plans (the σ = 1.5 grids of Ns = 128, 256,
// 512 and 1024); other lengths run the run-time plan.
#define SPEC_LAUNCH(len, kern, grid, nt, lds, ...)                          \
    do {                                                                    \
        switch (len) {                                                      \
            case 192: kern<CT, 192><<<(grid), (nt), (lds), s>>>(__VA_ARGS__); break;   \
            case 384: kern<CT, 384><<<(grid), (nt), (lds), s>>>(__VA_ARGS__); break;   \
            case 768: kern<CT, 768><<<(grid), (nt), (lds), s>>>(__VA_ARGS__); break;   \
            case 1536: kern<CT, 1536><<<(grid), (nt), (lds), s>>>(__VA_ARGS__); break; \
            default: kern<CT, 0><<<(grid), (nt), (lds), s>>>(__VA_ARGS__); break;      \
        }                                                                   \
    } while (0)

template <typename CT>
int forward_y_impl(vpbs_ctx *c) {
    hipStream_t s = c->s_lr;
    const SpecGeom g = spec_geom<CT>(c);
    const int64_t nkb = (c->nkx + (1 << g.lgB[1]) - 1) >> g.lgB[1];
    SPEC_LAUNCH(c->G[1], spec_y_forward, grid8(3 * (int64_t)c->G[2] * nkb), NTY, spec_lds<CT>(g, 1),
                g, (const C2<CT> *)c->ypen, (C2<CT> *)c->zpen, twiddles<CT>(c, 1));
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

template <typename CT>
int forward_impl(vpbs_ctx *c) {
    hipStream_t s = c->s_lr;
    const SpecGeom g = spec_geom<CT>(c);
    const int64_t nrows = 3 * (int64_t)c->G[2] * (c->G[1] / 2);
    SPEC_LAUNCH(c->G[0], spec_x_forward, grid8(nrows >> g.lgB[0]), NTX, spec_lds<CT>(g, 0),
                g, (const C2<CT> *)c->grid, (C2<CT> *)c->ypen, twiddles<CT>(c, 0));
    VPBS_HIP(c, hipGetLastError());
    return forward_y_impl<CT>(c);
}

// Slab kernels: stencil width W (2m), x-grid length L (compile-time plans).
template <typename CT, int W>
const void *slab_kernel_w(int L) {
    switch (L) {
        case 192: return (const void *)x_inverse_interp<CT, W, 192>;
        case 384: return (const void *)x_inverse_interp<CT, W, 384>;
        case 768: return (const void *)x_inverse_interp<CT, W, 768>;
        case 1536:  // FP32 only (one 156 KB block per CU; FP64 would need 312 KB)
            if constexpr (sizeof(CT) == 4) return (const void *)x_inverse_interp<CT, W, 1536>;
            return nullptr;
        default: return nullptr;
    }
}

template <typename CT>
const void *slab_kernel(int W, int L) {
    switch (W) {
        case 4: return slab_kernel_w<CT, 4>(L);
        case 6: return slab_kernel_w<CT, 6>(L);
        case 8: return slab_kernel_w<CT, 8>(L);
        case 10: return slab_kernel_w<CT, 10>(L);
        case 12: return slab_kernel_w<CT, 12>(L);
        case 14: return slab_kernel_w<CT, 14>(L);
        case 16: return slab_kernel_w<CT, 16>(L);
        default: return nullptr;
    }
}

template <typename CT>
size_t slab_lds(int L) {
    return sizeof(C2<CT>) * (size_t)L * (SLAB_BW + 1);  // tile and twiddles
}

template <typename CT>
int slab_interp_impl(vpbs_ctx *c, double *out) {
    if (!c->slab_interp) return fail(c, VPBS_ERR_STATE, "slab interpolation is not set up for this grid");
    const int W = 2 * c->p.nufft_m;
    const void *k = slab_kernel<CT>(W, c->G[0]);
    if (c->Np == 0) return VPBS_OK;
    SlabArgs a{};
    for (int d = 0; d < 3; d++) a.G[d] = c->G[d];
    a.nkx = c->nkx;
    a.Np = c->Np;
    a.start = c->nd_start;
    a.si0 = c->nd_i0;
    a.rec = c->nd_rec;
    a.part = c->slab_part;
    const C2<CT> *x1 = (const C2<CT> *)c->ypen;
    const C2<CT> *tw = twiddles<CT>(c, 0);
    void *args[] = {(void *)&a, (void *)&x1, (void *)&tw};
    VPBS_HIP(c, hipLaunchKernel(k, dim3(grid8(c->nslabs)), dim3(SLAB_NT), args, slab_lds<CT>(c->G[0]), c->s_lr));
    const int nslot = slab_nyh(W) * slab_nzh(W);
    slab_reduce<CT><<<(unsigned)((c->Np + 255) / 256), 256, 0, c->s_lr>>>(c->Np, nslot, (const CT *)c->slab_part,
                                                                          c->nd_perm, out);
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

template <typename CT>
int inverse_y_impl(vpbs_ctx *c, void *zp) {
    hipStream_t s = c->s_lr;
    const SpecGeom g = spec_geom<CT>(c);
    const int64_t nkb = (c->nkx + (1 << g.lgB[1]) - 1) >> g.lgB[1];
    SPEC_LAUNCH(c->G[1], spec_y_inverse, grid8(3 * (int64_t)c->G[2] * nkb), NTY, spec_lds<CT>(g, 1),
                g, (const C2<CT> *)zp, (C2<CT> *)c->ypen, twiddles<CT>(c, 1));
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

template <typename CT>
int fourier_impl(vpbs_ctx *c, uint32_t flags) {
    hipStream_t s = c->s_lr;
    const SpecGeom g = spec_geom<CT>(c);
    SpecFourier f = spec_fourier_args<CT>(c);
    f.do_vel = (flags & VPBS_VELOCITY) ? 1 : 0;
    f.do_psi = (flags & VPBS_STREAMFUNCTION) ? 1 : 0;
    f.keep_uhat = (flags & VPBS_KEEP_UHAT) ? 1 : 0;
    const int64_t nkb = (c->nkx + (1 << g.lgB[2]) - 1) >> g.lgB[2];
    C2<CT> *x2b = f.do_psi ? (C2<CT> *)c->zpen2 : nullptr;
    SPEC_LAUNCH(c->G[2], spec_z_fourier, grid8((int64_t)c->nky * nkb), NTZ, spec_lds<CT>(g, 2, 3),
                g, f, (C2<CT> *)c->zpen, x2b, twiddles<CT>(c, 2));
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

template <typename CT>
int inverse_impl(vpbs_ctx *c, void *zp) {
    hipStream_t s = c->s_lr;
    const SpecGeom g = spec_geom<CT>(c);
    const int64_t nrows = 3 * (int64_t)c->G[2] * (c->G[1] / 2);
    const int64_t nkb = (c->nkx + (1 << g.lgB[1]) - 1) >> g.lgB[1];
    SPEC_LAUNCH(c->G[1], spec_y_inverse, grid8(3 * (int64_t)c->G[2] * nkb), NTY, spec_lds<CT>(g, 1),
                g, (const C2<CT> *)zp, (C2<CT> *)c->ypen, twiddles<CT>(c, 1));
    SPEC_LAUNCH(c->G[0], spec_x_inverse, grid8(nrows >> g.lgB[0]), NTX, spec_lds<CT>(g, 0),
                g, (const C2<CT> *)c->ypen, (C2<CT> *)c->grid, twiddles<CT>(c, 0));
    VPBS_HIP(c, hipGetLastError());
    return VPBS_OK;
}

template <typename CT>
int band_impl(vpbs_ctx *c, double kmin2, double kmax2, double cf, double cc) {
    hipStream_t s = c->s_lr;
    const SpecGeom g = spec_geom<CT>(c);
    const SpecFourier f = spec_fourier_args<CT>(c);
    const int64_t nkb = (c->nkx + (1 << g.lgB[2]) - 1) >> g.lgB[2];
    SPEC_LAUNCH(c->G[2], spec_z_band, grid8((int64_t)c->nky * nkb), NTZ, spec_lds<CT>(g, 2),
                g, f, kmin2, kmax2, cf, cc, (C2<CT> *)c->zpen, twiddles<CT>(c, 2));
    VPBS_HIP(c, hipGetLastError());
    return inverse_impl<CT>(c, c->zpen);
}

template <typename CT>
int setup_impl(vpbs_ctx *c) {
    // twiddles e^{-2πin/G_d}, n ∈ [0, G_d), one table per axis
    std::vector<C2<CT>> tw;
    for (int d = 0; d < 3; d++) {
        const int L = c->G[d];
        for (int n = 0; n < L; n++) {
            const long double a = -2.0L * 3.14159265358979323846264338327950288L * n / L;
            tw.push_back(C2<CT>{(CT)std::cos(a), (CT)std::sin(a)});
        }
    }
    VPBS_HIP(c, hipMalloc(&c->spec_tw, sizeof(C2<CT>) * tw.size()));
    VPBS_HIP(c, hipMemcpy(c->spec_tw, tw.data(), sizeof(C2<CT>) * tw.size(), hipMemcpyHostToDevice));
    int lds_dev = 0;
    VPBS_HIP(c, hipDeviceGetAttribute(&lds_dev, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    c->spec_lds_max = std::min((size_t)lds_dev, (size_t)160 * 1024);
    const SpecGeom g = spec_geom<CT>(c);
    const int64_t nrows = 3 * (int64_t)c->G[2] * (c->G[1] / 2);
    if (nrows % ((int64_t)1 << g.lgB[0]))
        return fail(c, VPBS_ERR_INVALID, "fused transforms: %lld pair rows do not split into tiles of %d",
                    (long long)nrows, 1 << g.lgB[0]);
    const size_t lds_p = std::max(spec_lds<CT>(g, 0), spec_lds<CT>(g, 1)), lds_z = spec_lds<CT>(g, 2, 3);
    if (std::max(lds_p, lds_z) > (size_t)lds_dev)
        return fail(c, VPBS_ERR_INVALID, "fused transforms need %zu B of LDS per block, the device has %d",
                    std::max(lds_p, lds_z), lds_dev);
    auto attr = [&](const void *k, size_t lds) {
        return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    };
    auto attrs = [&](auto tag) -> int {
        constexpr int LS = decltype(tag)::value;
        VPBS_HIP(c, attr((const void *)spec_x_forward<CT, LS>, lds_p));
        VPBS_HIP(c, attr((const void *)spec_y_forward<CT, LS>, lds_p));
        VPBS_HIP(c, attr((const void *)spec_z_fourier<CT, LS>, lds_z));
        VPBS_HIP(c, attr((const void *)spec_z_band<CT, LS>, lds_z));
        VPBS_HIP(c, attr((const void *)spec_y_inverse<CT, LS>, lds_p));
        VPBS_HIP(c, attr((const void *)spec_x_inverse<CT, LS>, lds_p));
        return VPBS_OK;
    };
    VPBS_TRY(attrs(std::integral_constant<int, 0>{}));
    VPBS_TRY(attrs(std::integral_constant<int, 192>{}));
    VPBS_TRY(attrs(std::integral_constant<int, 384>{}));
    VPBS_TRY(attrs(std::integral_constant<int, 768>{}));
    VPBS_TRY(attrs(std::integral_constant<int, 1536>{}));
    if (spectral_slab_supported(c)) VPBS_HIP(c, attr(slab_kernel<CT>(2 * c->p.nufft_m, c->G[0]), slab_lds<CT>(c->G[0])));
    return VPBS_OK;
}

}  // namespace

bool spectral_supported(const vpbs_ctx *c) {
    const bool f32 = c->p.precision == VPBS_F32;
    for (int d = 0; d < 3; d++) {
        const AxisPlan a = plan_axis(c->G[d]);
        if (a.npass < 0 || c->G[d] > (f32 ? kind_tile<float>(d) : kind_tile<double>(d))) return false;
    }
    return true;
}

int spectral_setup(vpbs_ctx *c) {
    // separable factors (1/ψ̂)² e^{-k_d²/4α²} of the Fourier-space kernel
    const double inv4a2 = 1.0 / (4 * c->p.alpha * c->p.alpha);
    std::vector<double> hd((size_t)(c->kmax_mode[0] + 1) + (2 * c->kmax_mode[1] + 1) + (2 * c->kmax_mode[2] + 1));
    VPBS_HIP(c, hipMemcpy(hd.data(), c->deconv, sizeof(double) * hd.size(), hipMemcpyDeviceToHost));
    std::vector<double> fac(hd.size());
    size_t i = 0;
    for (int d = 0; d < 3; d++) {
        const int K = c->kmax_mode[d];
        const double ks = 2 * M_PI / c->p.Ls[d];
        for (int k = d == 0 ? 0 : -K; k <= K; k++, i++) {
            const double kk = ks * k;
            fac[i] = hd[i] * hd[i] * std::exp(-kk * kk * inv4a2);
        }
    }
    VPBS_HIP(c, hipMalloc((void **)&c->spec_fac, sizeof(double) * fac.size()));
    VPBS_HIP(c, hipMemcpy(c->spec_fac, fac.data(), sizeof(double) * fac.size(), hipMemcpyHostToDevice));
    return c->p.precision == VPBS_F32 ? setup_impl<float>(c) : setup_impl<double>(c);
}

int spectral_forward(vpbs_ctx *c) {
    return c->p.precision == VPBS_F32 ? forward_impl<float>(c) : forward_impl<double>(c);
}

bool spectral_slab_supported(const vpbs_ctx *c) {
    if (!c->fft_fused || c->p.fft != VPBS_FFT_FUSED) return false;
    const bool f32 = c->p.precision == VPBS_F32;
    const int W = 2 * c->p.nufft_m;
    if (!(f32 ? slab_kernel<float>(W, c->G[0]) : slab_kernel<double>(W, c->G[0]))) return false;
    // whole 8×8 blocks of slabs, and no home slab counted twice
    if (c->G[1] % (8 * SLAB_HY) || c->G[2] % (8 * SLAB_HZ)) return false;
    if (c->G[1] / SLAB_HY < slab_nyh(W) || c->G[2] / SLAB_HZ < slab_nzh(W)) return false;
    // FP32 on 768 cells: two blocks per CU (one loads x1 while the other transforms); FP64 on
    // 768 cells and FP32 on 1536: one, still faster than writing the real grid and gathering
    // from it (FP64, 768 cells: 7.1 against 8.1 ms at 1M nodes)
    return (f32 ? slab_lds<float>(c->G[0]) : slab_lds<double>(c->G[0])) <= c->spec_lds_max;
}

int spectral_slab_slots(const vpbs_ctx *c) { return slab_nyh(2 * c->p.nufft_m) * slab_nzh(2 * c->p.nufft_m); }

int spectral_y_inverse(vpbs_ctx *c, void *zp) {
    return c->p.precision == VPBS_F32 ? inverse_y_impl<float>(c, zp) : inverse_y_impl<double>(c, zp);
}

int spectral_x_inverse_interp(vpbs_ctx *c, double *out) {
    return c->p.precision == VPBS_F32 ? slab_interp_impl<float>(c, out) : slab_interp_impl<double>(c, out);
}



int spectral_fourier(vpbs_ctx *c, uint32_t flags) {
    return c->p.precision == VPBS_F32 ? fourier_impl<float>(c, flags) : fourier_impl<double>(c, flags);
}

int spectral_inverse(vpbs_ctx *c, void *zp) {
    return c->p.precision == VPBS_F32 ? inverse_impl<float>(c, zp) : inverse_impl<double>(c, zp);
}

int spectral_band(vpbs_ctx *c, double kmin2, double kmax2, double cf, double cc) {
    return c->p.precision == VPBS_F32 ? band_impl<float>(c, kmin2, kmax2, cf, cc)
                                      : band_impl<double>(c, kmin2, kmax2, cf, cc);
}

void spectral_release(vpbs_ctx *c) {
    if (c->spec_tw) (void)hipFree(c->spec_tw);
    if (c->spec_fac) (void)hipFree(c->spec_fac);
    c->spec_tw = nullptr;
    c->spec_fac = nullptr;
}

}  // namespace vpbs
