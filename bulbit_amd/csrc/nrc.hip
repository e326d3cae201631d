// nrc.hip — the radiance cache: a small MLP evaluated and trained with bf16
// MFMA on gfx950 (fp32 accumulation, fp32 master weights, Adam).
//
// Network (no biases): the 74-wide input encoding padded to kIn = 80,
// L hidden layers of width W with ReLU, and a linear output of 3 (stored as
// 16 columns, the last 13 held at zero):
//   h0 = enc(x);  h_{l+1} = relu(h_l W_l) for l < L;  y = h_L W_L
// Two shapes are built: W = 64, L = 4 and W = 128, L = 5.
// Input encoding of a raw row (16 floats: position in [0,1]^3, direction,
// normal, roughness, albedo): sin/cos(2^f pi p_a) frequency-major for
// f = 0..11 and axis a (72 values, the first 64 kept), then direction (3),
// normal (3), roughness (1), albedo (3), zeros to 80.
//
// Matrix products use v_mfma_f32_16x16x32_bf16: a 16x16 tile of A (rows x k)
// comes from lane l as A[l % 16][8 (l / 16) + j], j = 0..7; B (k x n) as
// B[8 (l / 16) + j][l % 16]; the f32 result D[4 (l / 16) + j][l % 16],
// j = 0..3 (a K of 16 left over takes the 16x16x16 form, 4 values per lane,
// whose result layout is the same).  The query kernel and the fused
// training step of the 64-wide network permute the k slots (frag_feature).
//
// Training step of the 128-wide network: the forward pass saves every
// layer's bf16 input per chunk of 64 rows, feature-major; k_mlp_dz runs the
// gradient chain dz_{l-1} = (dz_l W_l^T) masked by h_l > 0 per chunk and
// saves each dz_l, also feature-major; k_mlp_dw forms dW_l = h_l^T dz_l as a
// GEMM over the rows (whose operands are exactly those saved layouts), split
// over row ranges into partial sums; k_reduce_adam adds the partial sums in
// a fixed order and takes the Adam step.  The 64-wide network's step is one
// fused kernel (k_mlp_train).
#include "bb_math.h"
#include "../../include/bbgpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace bb
{
void set_error(const char* fmt, ...);
}

namespace bb
{

#ifndef BB_QTILES_C
#define BB_QTILES_C 2
#endif
#ifndef BB_QWAVES_C
#define BB_QWAVES_C 3
#endif
constexpr int kIn = 80;      // encoded input width (74 used)
constexpr int kOutCols = 16; // output columns (3 used)
constexpr int kRawStride = 16;
constexpr int kRows = 64;    // rows per wave per chunk (4 row tiles of 16)
constexpr int kNrcBlock = 256;

// One network shape.  Weight matrices: W_0 [kIn][W], W_1..W_{L-1} [W][W],
// W_L [W][kOutCols]; fp32 master copies k-major ([k][n]).
template <int W_, int L_>
struct Net
{
    static constexpr int W = W_, L = L_;
    static constexpr int k(int l) { return l == 0 ? kIn : W; }
    static constexpr int n(int l) { return l == L ? kOutCols : W; }
    static constexpr int off(int l) { return l == 0 ? 0 : off(l - 1) + k(l - 1) * n(l - 1); }
    static constexpr int params = off(L + 1);
    static constexpr int max_layer = (kIn * W > W * W ? kIn * W : W * W);
    // row stride (bf16) of a wave's activation tile in the forward pass
    static constexpr int ld = W > kIn ? W : kIn;
    // the whole bf16 network stays in LDS when it fits beside four
    // activation tiles; otherwise each layer is staged in turn
    static constexpr bool resident = params * 2 + 4 * kRows * ld * 2 <= 96 * 1024;
    // Layer inputs saved by the training forward pass, per chunk of kRows
    // rows, feature-major ([feature][row in chunk]): h_0 (kIn features), then
    // h_1..h_L (W each).  Eight consecutive rows of one feature are one 16-byte
    // MFMA operand (k = rows) of dW = h^T dz, and the ReLU mask of the dz
    // chain in its result layout.
    static constexpr int act_off(int l) { return l == 0 ? 0 : kRows * kIn + (l - 1) * kRows * W; }
    static constexpr int chunk_acts = act_off(L + 1);
    // dz_l (gradient of layer l's output, n(l) features) per chunk, feature-major
    static constexpr int dz_off(int l) { return l == 0 ? 0 : dz_off(l - 1) + n(l - 1) * kRows; }
    static constexpr int chunk_dz = dz_off(L + 1);
    // dW work groups (one wave each over a range of rows): layer 0 in
    // 80 x 64 blocks (5 x 4 tiles), the hidden layers in 64 x 64 blocks,
    // the output layer in 64 x 16 blocks
    static constexpr int groups = W / 64 + (L - 1) * (W / 64) * (W / 64) + W / 64;
    // The query kernel's weights: per layer, W_l^T cut into the A-operand
    // fragments of v_mfma_f32_16x16x32_bf16 (16 output features x 32 input
    // features, K padded to a multiple of 32), fragment (l, k-step, m-tile)
    // at frag_off(l) + ks * mt(l) + mt, 64 lanes x 8 bf16 each
    static constexpr int kpad(int l) { return (k(l) + 31) / 32 * 32; }
    static constexpr int mts(int l) { return n(l) / 16; }
    static constexpr int kss(int l) { return kpad(l) / 32; }
    static constexpr int frag_off(int l) { return l == 0 ? 0 : frag_off(l - 1) + mts(l - 1) * kss(l - 1); }
    static constexpr int frags = frag_off(L + 1);
    static constexpr int frag_elems = frags * 512;
    // query rows per wave and iteration (16 per row tile) and block size:
    // the wide network's fragments fill the LDS of a CU, so its blocks are
    // large; two row tiles share each weight fragment read from LDS
    static constexpr int qtiles = W > 64 ? 2 : BB_QTILES_C;
    static constexpr int qblock = W > 64 ? 512 : 256;
    static constexpr int qwaves = W > 64 ? 2 : BB_QWAVES_C; // waves per SIMD the query kernel is built for
    // The gradient chain's weights (fused training kernel): per layer
    // l >= 1, W_l cut into A-operand fragments of 16 input features of the
    // layer x 32 of its outputs (the output side padded to a multiple of 32,
    // same slot-to-feature map), fragment (l, k-step, m-tile) at
    // bfrag_off(l) + ks * bmts(l) + mt
    static constexpr int bmts(int l) { return k(l) / 16; }
    static constexpr int bkss(int l) { return (n(l) + 31) / 32; }
    static constexpr int bfrag_off(int l) { return l <= 1 ? 0 : bfrag_off(l - 1) + bmts(l - 1) * bkss(l - 1); }
    static constexpr int bfrags = bfrag_off(L + 1);
    static constexpr int bfrag_elems = bfrags * 512;
    // one fused kernel takes a training step of the narrow network (its
    // weights, in both fragment sets, fit in LDS beside the gradient tiles)
    static constexpr bool fused = W == 64;
};

using NetC = Net<64, 4>;  // the C4 cache
using NetL = Net<128, 5>; // the wide cache

// The rows a kernel works on.  total == nullptr: rows [0, cap).  Otherwise
// min(*total, cap) rows sit in the buffer (a count the device wrote) and,
// with batch > 0, one optimizer step takes batch number `step` of them,
// starting over when they run out: rows [off, off + len) with
// off = (step mod ceil(rows / batch)) batch.
struct RowSpan
{
    const int* total;
    int cap;
    int batch;
    int step;
};

BB_D void span_rows(const RowSpan& s, int& off, int& len)
{
    off = 0;
    len = s.cap;
    if (!s.total) return;
    const int rows = min(max(*s.total, 0), s.cap);
    len = rows;
    if (s.batch <= 0 || rows == 0) return;
    const int nb = (rows + s.batch - 1) / s.batch;
    off = (s.step % nb) * s.batch;
    len = min(s.batch, rows - off);
}

// most row ranges of the dW GEMM (partial sums held)
constexpr int kMaxSplits = 512;

// the fused training kernel (64-wide network): waves per block, rows per
// pass (16 per wave), row stride (bf16) of a feature's tile in LDS (272
// bytes), most blocks (one per CU: its LDS)
constexpr int kTrainWaves = 8;
constexpr int kTrainRows = 16 * kTrainWaves;
constexpr int kTrainLd = kTrainRows + 8;
constexpr int kTrainBlocks = 256;
static_assert(kTrainBlocks <= kMaxSplits, "one partial sum per block");

// row ranges for the dW GEMM: about 2048 waves in all, at least four chunks each
BB_HD int dw_splits(int chunks, int groups)
{
    const int want = (2048 + groups - 1) / groups;
    const int a = want < chunks / 4 ? want : chunks / 4;
    const int b = a < kMaxSplits ? a : kMaxSplits;
    return b > 1 ? b : 1;
}

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (finite inputs)
BB_HD uint16_t to_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

BB_D float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));

// two floats as bf16 (round to nearest even: one v_cvt_pk_bf16_f32), a in
// the low half; the same bits as to_bf16 for finite values
BB_D uint32_t pack_bf16(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2v{ a, b }, bf16x2));
}

// --------------------------------------------------------------- encoding

// sin and cos of pi v 2^f, as a packed bf16 pair (sin low), for
// half_scale = 2^(f - 1).  The angle in turns, t = v 2^(f - 1), less its
// nearest integer (both steps exact) lies in [-1/2, 1/2], inside the range
// of the hardware's v_sin_f32 / v_cos_f32 (sin and cos of 2 pi t).
BB_D uint32_t sincos_pi_bf16(float v, float half_scale)
{
    float t = v * half_scale;
    t = t - __builtin_rintf(t);
    return pack_bf16(__builtin_amdgcn_sinf(t), __builtin_amdgcn_cosf(t));
}

// the kIn encoded features of one raw row (out 4-byte aligned): pair p < 32
// is sin/cos(2^(p / 3) pi r[p mod 3]), features 64..73 are r[3..12]
BB_D void encode_row(const float* r, uint16_t* out)
{
    uint32_t* o = reinterpret_cast<uint32_t*>(out);
#pragma unroll
    for (int p = 0; p < 32; ++p) o[p] = sincos_pi_bf16(r[p % 3], 0.5f * float(1 << (p / 3)));
#pragma unroll
    for (int i = 0; i < 5; ++i) o[32 + i] = pack_bf16(r[3 + 2 * i], r[4 + 2 * i]);
    static_assert(kIn == 80, "features 74..79 are the padding");
    o[37] = o[38] = o[39] = 0u;
}

// ------------------------------------------------------------ MFMA pieces

BB_D s4 load4(const uint16_t* p) { return *reinterpret_cast<const s4*>(p); }

BB_D f4v mfma(s4 a, s4 b, f4v c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s8v __attribute__((ext_vector_type(8)));

BB_D s8v load8(const uint16_t* p) { return *reinterpret_cast<const s8v*>(p); }

// v_mfma_f32_16x16x32_bf16: K = 32 per instruction, the same result layout
// as the 16x16x16 form; lane l holds k = 8 (l / 16) .. + 7 of its A row and
// B column (in the plain-layout kernels)
BB_D f4v mfma32(s8v a, s8v b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                  0, 0);
}

// one layer for a wave's kRows rows: act [kRows][ld_in] (bf16, LDS) times
// W_l (WT: [n][k] bf16, LDS) -> acc[row tile][col tile]
template <int K, int N>
BB_D void layer_forward(const uint16_t* act, int ld_in, const uint16_t* WT, f4v (&acc)[4][N / 16], int lane)
{
    const int r16 = lane & 15, q = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < N / 16; ++ct) acc[rt][ct] = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int kk = 0; kk < K / 32; ++kk) {
        s8v b[N / 16];
#pragma unroll
        for (int ct = 0; ct < N / 16; ++ct) b[ct] = load8(WT + (ct * 16 + r16) * K + kk * 32 + q * 8);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const s8v a = load8(act + (rt * 16 + r16) * ld_in + kk * 32 + q * 8);
#pragma unroll
            for (int ct = 0; ct < N / 16; ++ct) acc[rt][ct] = mfma32(a, b[ct], acc[rt][ct]);
        }
    }
    if constexpr (K % 32 == 16) { // the last 16 of K (layer 0: K = 80) in the 16x16x16 form
        constexpr int k0 = K - 16;
        s4 b[N / 16];
#pragma unroll
        for (int ct = 0; ct < N / 16; ++ct) b[ct] = load4(WT + (ct * 16 + r16) * K + k0 + q * 4);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const s4 a = load4(act + (rt * 16 + r16) * ld_in + k0 + q * 4);
#pragma unroll
            for (int ct = 0; ct < N / 16; ++ct) acc[rt][ct] = mfma(a, b[ct], acc[rt][ct]);
        }
    }
    static_assert(K % 16 == 0, "K in steps of 16");
}

// bf16 copy of count values (a multiple of 8) into LDS by the whole block
BB_D void stage16(uint16_t* dst, const uint16_t* src, int count)
{
    for (int i = threadIdx.x; i < count / 8; i += blockDim.x)
        reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
}

// ---------------------------------------------------------------- forward

struct NrcWeights
{
    const uint16_t* WT; // bf16 [n][k] per layer (at off(l))
    const uint16_t* W;  // bf16 [k][n] per layer
};

// y = f(enc(raw)) for n rows; with SAVE the bf16 input of every layer is
// stored per chunk, feature-major (Net::act_off), for the backward pass.
// The block's four waves take four consecutive chunks per iteration; when
// the weights are staged layer by layer every wave joins every barrier.
template <class NT, bool SAVE>
__global__ __launch_bounds__(kNrcBlock) void k_mlp_forward(const float* __restrict__ raw, RowSpan span, NrcWeights Wt,
                                                            float* __restrict__ out, uint16_t* __restrict__ acts)
{
    constexpr int W = NT::W, L = NT::L, LD = NT::ld;
    int off = 0, n = 0;
    span_rows(span, off, n);
    raw += size_t(off) * kRawStride;
    __shared__ __align__(16) uint16_t sW[NT::resident ? NT::params : NT::max_layer];
    __shared__ __align__(16) uint16_t sAct[kNrcBlock / 64][kRows * LD];
    static_assert(NT::params % 8 == 0 && NT::max_layer % 8 == 0, "weights staged 16 bytes at a time");
    if constexpr (NT::resident) {
        stage16(sW, Wt.WT, NT::params);
        __syncthreads();
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* act = sAct[wave];
    const int r16 = lane & 15, q = lane >> 4;
    const int chunks = (n + kRows - 1) / kRows;
    for (int base = blockIdx.x * (kNrcBlock / 64); base < chunks; base += gridDim.x * (kNrcBlock / 64)) {
        const int ch = base + wave;
        const bool active = ch < chunks;
        const int row0 = ch * kRows;
        if (active) {
            // each lane encodes its row straight into the wave's LDS rows
            const int row = row0 + lane;
            uint16_t* e = act + lane * LD;
            if (row < n)
                encode_row(raw + size_t(row) * kRawStride, e);
            else
                for (int i = 0; i < kIn; ++i) e[i] = 0;
            if (SAVE) // every row of the chunk, zeros beyond n
                for (int i = 0; i < kIn; ++i) acts[size_t(ch) * NT::chunk_acts + NT::act_off(0) + i * kRows + lane] = e[i];
        }
        // hidden layers in a rolled loop (one copy of the code: small kernels
        // stay in the instruction cache)
#pragma unroll 1
        for (int l = 0; l < L; ++l) {
            const uint16_t* Wl = sW + (NT::resident ? NT::off(l) : 0);
            if constexpr (!NT::resident) {
                __syncthreads(); // every wave is done with the previous layer's weights
                stage16(sW, Wt.WT + NT::off(l), NT::k(l) * NT::n(l));
                __syncthreads();
            }
            if (!active) continue;
            f4v acc[4][W / 16];
            if (l == 0)
                layer_forward<kIn, W>(act, LD, Wl, acc, lane);
            else
                layer_forward<W, W>(act, LD, Wl, acc, lane);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                for (int ct = 0; ct < W / 16; ++ct) {
                    s4 hv;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int r = rt * 16 + q * 4 + j, c = ct * 16 + r16;
                        const uint16_t h = to_bf16(fmaxf(acc[rt][ct][j], 0.0f));
                        act[r * LD + c] = h;
                        hv[j] = short(h);
                    }
                    // rows rt*16 + 4q .. +3 of feature ct*16 + r16 of h_{l+1}
                    if (SAVE)
                        *reinterpret_cast<s4*>(acts + size_t(ch) * NT::chunk_acts + NT::act_off(l + 1) +
                                               (ct * 16 + r16) * kRows + rt * 16 + q * 4) = hv;
                }
        }
        // output layer
        const uint16_t* Wo = sW + (NT::resident ? NT::off(L) : 0);
        if constexpr (!NT::resident) {
            __syncthreads();
            stage16(sW, Wt.WT + NT::off(L), NT::k(L) * NT::n(L));
            __syncthreads();
        }
        if (!active) continue;
        f4v acc[4][1];
        layer_forward<W, kOutCols>(act, LD, Wo, acc, lane);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = rt * 16 + q * 4 + j;
                if (r16 < 3 && row0 + r < n) out[size_t(row0 + r) * 3 + r16] = acc[rt][0][j];
            }
    }
}

// --------------------------------------------------------------- backward

// relative L2 loss of NRC: mean over rows and channels of
// (y - t)^2 / (sg(y)^2 + 0.01); returns dL/dy
BB_D float loss_grad(float y, float t, float inv_count, float& loss)
{
    const float d = y - t;
    const float den = y * y + 0.01f;
    loss += d * d / den * inv_count;
    return 2.0f * d / den * inv_count;
}

// dz_{l-1} = (dz_l W_l^T) masked by h_l > 0 for a wave's chunk: dz_l
// (kRows x N) row-major in dzl (LDS), W_l k-major from global memory, h_l
// (the mask) feature-major from the saved inputs; the result replaces dzl
// (row-major, stride W) and is stored feature-major to dzc.
template <class NT, int N>
BB_D void dz_layer(int l, const uint16_t* hc, const uint16_t* Wk, uint16_t* dzl, uint16_t* dzc, int lane)
{
    constexpr int W = NT::W;
    const int r16 = lane & 15, q = lane >> 4;
    const uint16_t* h = hc + NT::act_off(l);
    const uint16_t* Wl = Wk + NT::off(l);
    // the mask first, so that its loads overlap the products
    s4 hv[4][W / 16];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < W / 16; ++ct) hv[rt][ct] = load4(h + (ct * 16 + r16) * kRows + rt * 16 + q * 4);
    f4v acc[4][W / 16];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < W / 16; ++ct) acc[rt][ct] = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
    // (rows x N) (N x W): A = dz rows, B = W_l^T (B[n][k] = W_l[k][n])
#pragma unroll
    for (int kk = 0; kk < N / 32; ++kk) {
        s8v b[W / 16];
#pragma unroll
        for (int ct = 0; ct < W / 16; ++ct) b[ct] = load8(Wl + (ct * 16 + r16) * N + kk * 32 + q * 8);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const s8v av = load8(dzl + (rt * 16 + r16) * W + kk * 32 + q * 8);
#pragma unroll
            for (int ct = 0; ct < W / 16; ++ct) acc[rt][ct] = mfma32(av, b[ct], acc[rt][ct]);
        }
    }
    if constexpr (N % 32 == 16) { // the last 16 of K (the output layer: N = 16) in the 16x16x16 form
        constexpr int k0 = N - 16;
        s4 b[W / 16];
#pragma unroll
        for (int ct = 0; ct < W / 16; ++ct) b[ct] = load4(Wl + (ct * 16 + r16) * N + k0 + q * 4);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const s4 av = load4(dzl + (rt * 16 + r16) * W + k0 + q * 4);
#pragma unroll
            for (int ct = 0; ct < W / 16; ++ct) acc[rt][ct] = mfma(av, b[ct], acc[rt][ct]);
        }
    }
    uint16_t* out = dzc + NT::dz_off(l - 1);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < W / 16; ++ct) {
            s4 dv;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint16_t hj = uint16_t(hv[rt][ct][j]);
                const bool on = hj != 0 && !(hj & 0x8000u);
                const uint16_t d = on ? to_bf16(acc[rt][ct][j]) : uint16_t(0);
                dv[j] = short(d);
                dzl[(rt * 16 + q * 4 + j) * W + ct * 16 + r16] = d;
            }
            *reinterpret_cast<s4*>(out + (ct * 16 + r16) * kRows + rt * 16 + q * 4) = dv;
        }
}

// The gradient chain for n rows whose layer inputs the forward pass saved:
// dz_L = dL/dy, then dz_{l-1} for l = L..1, every dz_l stored per chunk
// (Net::dz_off).  One wave per chunk; each block leaves its loss in
// loss_part[block] (zero without rows).
template <class NT>
__global__ __launch_bounds__(kNrcBlock) void k_mlp_dz(const float* __restrict__ y, const float* __restrict__ target,
                                                      RowSpan span, const uint16_t* __restrict__ Wk,
                                                      const uint16_t* __restrict__ acts, uint16_t* __restrict__ dzs,
                                                      float* __restrict__ loss_part)
{
    constexpr int W = NT::W, L = NT::L;
    int off = 0, n = 0;
    span_rows(span, off, n);
    target += size_t(off) * 3;
    __shared__ __align__(16) uint16_t sD[kNrcBlock / 64][kRows * W]; // dz of the current layer output, row-major
    __shared__ float sLoss[kNrcBlock / 64];
    static_assert(kNrcBlock / 64 == 4, "four wave sums");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, q = lane >> 4;
    uint16_t* dzl = sD[wave];
    const float inv_count = n > 0 ? 1.0f / (3.0f * float(n)) : 0.0f;
    float loss = 0.0f;
    const int chunks = (n + kRows - 1) / kRows;
    for (int ch = blockIdx.x * (kNrcBlock / 64) + wave; ch < chunks; ch += gridDim.x * (kNrcBlock / 64)) {
        const int row0 = ch * kRows;
        const uint16_t* hc = acts + size_t(ch) * NT::chunk_acts;
        uint16_t* dzc = dzs + size_t(ch) * NT::chunk_dz;
        // dz_L: dL/dy of rows rt*16 + 4q + j, column r16 (3 used)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            s4 dv;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = rt * 16 + q * 4 + j, row = row0 + r;
                float g = 0.0f;
                if (r16 < 3 && row < n)
                    g = loss_grad(y[size_t(row) * 3 + r16], target[size_t(row) * 3 + r16], inv_count, loss);
                const uint16_t gb = to_bf16(g);
                dv[j] = short(gb);
                dzl[r * W + r16] = gb;
            }
            *reinterpret_cast<s4*>(dzc + NT::dz_off(L) + r16 * kRows + rt * 16 + q * 4) = dv;
        }
        dz_layer<NT, kOutCols>(L, hc, Wk, dzl, dzc, lane);
#pragma unroll 1
        for (int l = L - 1; l >= 1; --l) dz_layer<NT, W>(l, hc, Wk, dzl, dzc, lane);
    }
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
    if (lane == 0) sLoss[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = ((sLoss[0] + sLoss[1]) + sLoss[2]) + sLoss[3];
}

// dW_l rows [mt0, mt0 + 16 TM) x columns [nt0, nt0 + 16 TN) summed over
// chunks [c0, c1): A = h_l^T (features x rows), B = dz_l (rows x features),
// both straight from their feature-major chunks; out is the layer's block of
// one partial sum ([K][N])
template <class NT, int TM, int TN>
BB_D void dw_tile(int l, int mt0, int nt0, int c0, int c1, const uint16_t* acts, const uint16_t* dzs, float* out,
                  int lane)
{
    const int r16 = lane & 15, q = lane >> 4;
    const int N = NT::n(l);
    f4v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
    for (int c = c0; c < c1; ++c) {
        const uint16_t* h = acts + size_t(c) * NT::chunk_acts + NT::act_off(l) + (mt0 * 16 + r16) * kRows + q * 8;
        const uint16_t* d = dzs + size_t(c) * NT::chunk_dz + NT::dz_off(l) + (nt0 * 16 + r16) * kRows + q * 8;
#pragma unroll
        for (int st = 0; st < kRows / 32; ++st) {
            s8v a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = load8(h + i * 16 * kRows + st * 32);
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = load8(d + j * 16 * kRows + st * 32);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = mfma32(a[i], b[j], acc[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                out[((mt0 + i) * 16 + q * 4 + e) * N + (nt0 + j) * 16 + r16] = acc[i][j][e];
}

// dW of every layer as Net::groups work groups times `splits` row ranges,
// one wave each: block b does group b % groups over chunks
// [chunks s / splits, chunks (s + 1) / splits) with s = b / groups, into
// partials[s][params]
template <class NT>
__global__ __launch_bounds__(64) void k_mlp_dw(const uint16_t* __restrict__ acts, const uint16_t* __restrict__ dzs,
                                               RowSpan span, float* __restrict__ partials)
{
    constexpr int W = NT::W, L = NT::L, B = W / 64; // 64-wide blocks per hidden dimension
    int off = 0, n = 0;
    span_rows(span, off, n);
    const int chunks = (n + kRows - 1) / kRows;
    const int splits = dw_splits(chunks, NT::groups);
    const int g = blockIdx.x % NT::groups, s = blockIdx.x / NT::groups;
    if (n == 0 || s >= splits) return;
    const int c0 = int(int64_t(chunks) * s / splits), c1 = int(int64_t(chunks) * (s + 1) / splits);
    const int lane = threadIdx.x;
    float* out = partials + size_t(s) * NT::params;
    if (g < B) { // layer 0: all 80 rows, columns 64 g ..
        dw_tile<NT, 5, 4>(0, 0, 4 * g, c0, c1, acts, dzs, out + NT::off(0), lane);
    } else if (g < B + (L - 1) * B * B) {
        const int i = g - B, l = 1 + i / (B * B), blk = i % (B * B);
        dw_tile<NT, 4, 4>(l, 4 * (blk / B), 4 * (blk % B), c0, c1, acts, dzs, out + NT::off(l), lane);
    } else { // output layer: rows 64 (g - ..) .., its 16 columns
        dw_tile<NT, 4, 1>(L, 4 * (g - B - (L - 1) * B * B), 0, c0, c1, acts, dzs, out + NT::off(L), lane);
    }
}

// k-step, lane group and slot of input feature k in the 16x16x32 operands
// (the inverse of frag_feature)
BB_HD void frag_slot(int k, int& ks, int& g, int& j)
{
    ks = k / 32;
    const int r = k % 32;
    g = (r % 16) / 4;
    j = r % 4 + (r < 16 ? 0 : 4);
}

// The gradient as the sum of the row ranges' partial sums, kept in g: wave
// w of a block adds ranges w, w + 4, ... in order, then the four sums are
// added in wave order.  Then Adam (beta1 0.9, beta2 0.99, eps 1e-8) on the
// fp32 master weights when apply, and the bf16 copies: both matrix layouts,
// the query's fragments and (fused network) the gradient chain's; the unused
// output columns stay zero.  reduce = 0 only refreshes the copies.  The
// loss is the sum of the step's per-block partial sums (nloss of them) in a
// fixed order.  fused_grid > 0: the partial sums came from that many blocks
// of the fused kernel, one per pass while passes last.
template <class NT>
__global__ __launch_bounds__(256) void k_reduce_adam(float* __restrict__ w, float* __restrict__ m,
                                                     float* __restrict__ v, float* __restrict__ g,
                                                     const float* __restrict__ partials, RowSpan span, int reduce,
                                                     int fused_grid, float lr, float c1, float c2, int apply,
                                                     uint16_t* __restrict__ WT, uint16_t* __restrict__ W,
                                                     uint16_t* __restrict__ WF, uint16_t* __restrict__ WB,
                                                     const float* __restrict__ loss_part, int nloss,
                                                     float* __restrict__ loss, int* __restrict__ diverged)
{
    __shared__ float part[4][64];
    int off = 0, n = 0;
    span_rows(span, off, n);
    // no rows: nothing to reduce and no step
    int splits = 0;
    if (reduce && n > 0) {
        const int passes = (n + kTrainRows - 1) / kTrainRows;
        splits = fused_grid > 0 ? min(fused_grid, passes) : dw_splits((n + kRows - 1) / kRows, NT::groups);
    }
    if (reduce && n == 0) apply = 0;
    const int t = threadIdx.x, wv = t >> 6;
    const int i = blockIdx.x * 64 + (t & 63);
    float s = 0.0f;
    if (i < NT::params) {
#pragma unroll 8
        for (int b = wv; b < splits; b += 4) s += partials[size_t(b) * NT::params + i];
    }
    part[wv][t & 63] = s;
    __syncthreads();
    if (wv != 0) return;
    float ls = 0.0f;
    if (reduce) {
        for (int b = t; b < nloss; b += 64) ls += loss_part[b];
        for (int o = 32; o > 0; o >>= 1) ls += __shfl_xor(ls, o);
        if (blockIdx.x == 0 && t == 0) *loss = ls;
    }
    if (i >= NT::params) return;
    const float gi = ((part[0][t] + part[1][t]) + part[2][t]) + part[3][t];
    int l = 0;
    while (l < NT::L && i >= NT::off(l + 1)) ++l;
    const int K = NT::k(l), N = NT::n(l);
    const int k = (i - NT::off(l)) / N, nn = (i - NT::off(l)) % N;
    if (splits > 0) g[i] = gi;
    float wi = w[i];
    if (apply && !(l == NT::L && nn >= 3)) {
        const float mi = 0.9f * m[i] + 0.1f * gi;
        const float vi = 0.99f * v[i] + 0.01f * gi * gi;
        m[i] = mi;
        v[i] = vi;
        wi -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
        w[i] = wi;
    }
    // a diverged network (non-finite loss or weight) is flagged for good
    if (!isfinite(wi) || (apply && !isfinite(ls))) *diverged = 1;
    const uint16_t h = to_bf16(wi);
    W[i] = h;
    WT[NT::off(l) + nn * K + k] = h;
    // query fragment (l, ks, nn / 16): W_l^T, output feature nn on the lane,
    // input feature k in the slot
    int ks, gg, j;
    frag_slot(k, ks, gg, j);
    WF[(NT::frag_off(l) + ks * NT::mts(l) + nn / 16) * 512 + (16 * gg + nn % 16) * 8 + j] = h;
    if constexpr (NT::fused) {
        // gradient-chain fragment (l, ks, k / 16): W_l, input feature k on
        // the lane, output feature nn in the slot
        if (l >= 1) {
            frag_slot(nn, ks, gg, j);
            WB[(NT::bfrag_off(l) + ks * NT::bmts(l) + k / 16) * 512 + (16 * gg + k % 16) * 8 + j] = h;
        }
    }
}

// ------------------------------------------------ query (16x16x32 MFMA)
//
// v_mfma_f32_16x16x32_bf16: lane l holds 8 bf16 of A (row l % 16) and 8 of B
// (column l % 16) for the same 8 values of k, chosen by the lane group
// g = l / 16 and the slot j; the result D[4 g + j][l % 16], j = 0..3.  The
// network runs transposed, y^T = W^T h^T: a layer's output tile (16
// features x 16 rows) sits in D as features 4 g + j of row l % 16, so two
// consecutive output tiles, relu'd and rounded to bf16, are the next layer's
// B operand for one k-step when slot j < 4 means feature 32 ks + 4 g + j and
// slot j >= 4 feature 32 ks + 16 + 4 g + (j - 4).  The weight fragments use
// the same slot-to-feature map, so activations never leave the registers.


// the input feature of k-step ks, lane group g, slot j
BB_HD constexpr int frag_feature(int ks, int g, int j) { return 32 * ks + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4)); }

// one layer for R row tiles: acc[mt][r] = sum over k-steps of A(l, ks, mt) B[ks][r]
template <int KS, int MT, int R>
BB_D void qlayer(const uint16_t* frags, const s8v (&B)[4][R], f4v (&acc)[MT][R], int lane)
{
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < R; ++r) acc[mt][r] = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const s8v a = *reinterpret_cast<const s8v*>(frags + (ks * MT + mt) * 512 + lane * 8);
#pragma unroll
            for (int r = 0; r < R; ++r) acc[mt][r] = mfma32(a, B[ks][r], acc[mt][r]);
        }
}

// relu of a packed bf16 pair: as 16-bit integers the negative values (sign
// bit set) are the negative ones, so max(x, 0) per half keeps x >= 0 and
// zeroes the rest (-0 included)
BB_D uint32_t relu_bf16x2(uint32_t x)
{
    typedef short s2v __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s2v, x), s2v{ 0, 0 }));
}

// relu, bf16, and the output tiles paired into the next layer's B operands
template <int MT, int R>
BB_D void qactivate(const f4v (&acc)[MT][R], s8v (&B)[4][R])
{
#pragma unroll
    for (int ks = 0; ks < MT / 2; ++ks)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const f4v lo = acc[2 * ks][r], hi = acc[2 * ks + 1][r];
            const u4v d{ relu_bf16x2(pack_bf16(lo[0], lo[1])), relu_bf16x2(pack_bf16(lo[2], lo[3])),
                         relu_bf16x2(pack_bf16(hi[0], hi[1])), relu_bf16x2(pack_bf16(hi[2], hi[3])) };
            B[ks][r] = __builtin_bit_cast(s8v, d);
        }
}

// The lane's eight sin/cos pairs of k-steps 0 and 1 of the layer-0 operand:
// slots 2i, 2i + 1 of k-step ks = i / 4 are features 2p, 2p + 1 with
// p = 16 ks + 8 ((i / 2) mod 2) + 2 g + i mod 2 (frag_feature's map), the
// sin/cos of 2^(p / 3) pi times axis p mod 3
struct EncLane
{
    float scale[8];
    int axis[8];
};

BB_D EncLane enc_lane(int g)
{
    static_assert(frag_feature(1, 3, 7) == 63, "k-steps 0 and 1 hold the 64 sin/cos features");
    EncLane e;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int p = 16 * (i >> 2) + 8 * ((i >> 1) & 1) + 2 * g + (i & 1);
        e.scale[i] = 0.5f * float(1 << (p / 3));
        e.axis[i] = p % 3;
    }
    return e;
}

// the layer-0 B operands (k-steps 0..2) of one row: the lane's 24 features
// of the raw row rp (zeros when !ok)
BB_D void enc_operands(const float* rp, bool ok, int g, const EncLane& el, s8v& b0, s8v& b1, s8v& b2)
{
    float xyz[3] = { 0.0f, 0.0f, 0.0f };
    if (ok) {
        xyz[0] = rp[0];
        xyz[1] = rp[1];
        xyz[2] = rp[2];
    }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        u4v d;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int a = el.axis[4 * ks + i];
            const float v = a == 0 ? xyz[0] : (a == 1 ? xyz[1] : xyz[2]);
            d[i] = sincos_pi_bf16(v, el.scale[4 * ks + i]);
        }
        (ks == 0 ? b0 : b1) = __builtin_bit_cast(s8v, d);
    }
    // k-step 2: slots 0..3 are features 64 + 4 g + j = r[3 + 4 g + j] (below
    // 74), slots 4..7 features 80.. (padding)
    float e[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) e[j] = (ok && 4 * g + j < 10) ? rp[3 + 4 * g + j] : 0.0f;
    const u4v d{ pack_bf16(e[0], e[1]), pack_bf16(e[2], e[3]), 0u, 0u };
    b2 = __builtin_bit_cast(s8v, d);
}

// y = f(enc(raw)) for the rows span selects, weights as fragments in LDS
template <class NT>
__global__ __launch_bounds__(NT::qblock, NT::qwaves) void k_mlp_query(const float* __restrict__ raw, RowSpan span,
                                                         const uint16_t* __restrict__ wf, float* __restrict__ out)
{
    constexpr int W = NT::W, L = NT::L, R = NT::qtiles;
    __shared__ __align__(16) uint16_t sWF[NT::frag_elems];
    stage16(sWF, wf, NT::frag_elems);
    __syncthreads();
    int off = 0, n = 0;
    span_rows(span, off, n);
    raw += size_t(off) * kRawStride;
    const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
    const int waves = blockDim.x >> 6;
    constexpr int rows = 16 * R;
    static_assert(NT::kss(0) == 3, "layer 0 takes three k-steps");
    const EncLane el = enc_lane(g);
    for (int row0 = (blockIdx.x * waves + int(threadIdx.x >> 6)) * rows; row0 < n;
         row0 += gridDim.x * waves * rows) {
        s8v B[4][R];
        // layer 0 input: each lane encodes 24 features of row c16 of each
        // tile (rows beyond n encode zeros; their outputs are not stored)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = row0 + r * 16 + c16;
            const bool ok = row < n;
            enc_operands(raw + size_t(ok ? row : 0) * kRawStride, ok, g, el, B[0][r], B[1][r], B[2][r]);
        }
        {
            f4v acc[NT::mts(0)][R];
            qlayer<NT::kss(0), NT::mts(0), R>(sWF + NT::frag_off(0) * 512, B, acc, lane);
            qactivate<NT::mts(0), R>(acc, B);
        }
#pragma unroll 1
        for (int l = 1; l < L; ++l) {
            f4v acc[W / 16][R];
            qlayer<W / 32, W / 16, R>(sWF + NT::frag_off(l) * 512, B, acc, lane);
            qactivate<W / 16, R>(acc, B);
        }
        f4v acc[1][R];
        qlayer<W / 32, 1, R>(sWF + NT::frag_off(L) * 512, B, acc, lane);
        // channels 0..2 are rows 0..2 of the output tile: lanes of group 0
        if (g == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int row = row0 + r * 16 + c16;
                if (row < n) {
                    out[size_t(row) * 3] = acc[0][r][0];
                    out[size_t(row) * 3 + 1] = acc[0][r][1];
                    out[size_t(row) * 3 + 2] = acc[0][r][2];
                }
            }
        }
    }
}

// ------------------------------------- fused training step (64-wide network)
//
// One block of kTrainWaves waves takes kTrainRows rows per pass, 16 per
// wave: the forward pass as in the query kernel (transposed, activations in
// registers), keeping every layer's input operand; dL/dy; then the gradient
// chain dz_{l-1}^T = W_l dz_l^T masked by h_l > 0, also transposed, with the
// gradient-chain fragments in LDS.  For dW_l = h_l^T dz_l, whose MFMA needs
// the rows along k, each wave writes h_l and dz_l feature-major
// ([feature][row]) into LDS; after a barrier the block's waves add the
// products of their share of the 16x16 dW tiles over the pass's rows to
// accumulators held across passes.  The tiles are double-buffered, so one
// barrier per layer suffices: a buffer is written again only two layers
// later, after every wave has passed the barrier that follows the reads.
// Each block leaves one partial sum of dW and of the loss.


// most blocks of the fused training kernel for max_len rows
BB_HD int train_grid(int max_len)
{
    const int passes = (max_len + kTrainRows - 1) / kTrainRows;
    return passes < 1 ? 1 : (passes < kTrainBlocks ? passes : kTrainBlocks);
}

// bf16 pair x where the matching bf16 of h is positive (as 16-bit integers:
// nonzero with the sign bit clear), zero elsewhere
BB_D uint32_t mask_positive(uint32_t x, uint32_t h)
{
    typedef short s2v __attribute__((ext_vector_type(2)));
    const s2v on = __builtin_bit_cast(s2v, h) > s2v{ 0, 0 };
    return x & __builtin_bit_cast(uint32_t, on);
}

// a wave's 16 rows of operand B (slot j of k-step ks is feature
// frag_feature(ks, g, j)) into the feature-major tile s, features below NF
template <int KS, int NF>
BB_D void store_features(uint16_t* s, const s8v (&B)[4][1], int wave, int lane)
{
    const int g = lane >> 4, c16 = lane & 15;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int f = frag_feature(ks, g, j);
            if (f < NF) s[f * kTrainLd + 16 * wave + c16] = uint16_t(B[ks][0][j]);
        }
}

// dW tiles t = wave + kTrainWaves i of an MT x NT grid (t = mt NT + nt):
// acc[i] += h^T dz over the pass's rows, both from feature-major tiles
template <int MT, int NT, int NI>
BB_D void dw_pass(const uint16_t* sh, const uint16_t* sz, f4v (&acc)[NI], int wave, int lane)
{
    const int g = lane >> 4, c16 = lane & 15;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int t = wave + kTrainWaves * i;
        if (t >= MT * NT) continue;
        const int mt = t / NT, nt = t % NT;
        const uint16_t* a = sh + (16 * mt + c16) * kTrainLd + 8 * g;
        const uint16_t* b = sz + (16 * nt + c16) * kTrainLd + 8 * g;
#pragma unroll
        for (int ks = 0; ks < kTrainRows / 32; ++ks)
            acc[i] = mfma32(*reinterpret_cast<const s8v*>(a + 32 * ks), *reinterpret_cast<const s8v*>(b + 32 * ks),
                            acc[i]);
    }
}

// the accumulated dW tiles of one layer into a partial sum ([k][n], k-major)
template <int MT, int NT, int NI>
BB_D void dw_store(float* out, int N, const f4v (&acc)[NI], int wave, int lane)
{
    const int g = lane >> 4, c16 = lane & 15;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int t = wave + kTrainWaves * i;
        if (t >= MT * NT) continue;
        const int mt = t / NT, nt = t % NT;
#pragma unroll
        for (int j = 0; j < 4; ++j) out[(16 * mt + 4 * g + j) * N + 16 * nt + c16] = acc[i][j];
    }
}

constexpr int tiles_per_wave(int tiles) { return (tiles + kTrainWaves - 1) / kTrainWaves; }

template <class NT>
__global__ __launch_bounds__(64 * kTrainWaves) void k_mlp_train(const float* __restrict__ raw,
                                                                const float* __restrict__ target, RowSpan span,
                                                                const uint16_t* __restrict__ wf,
                                                                const uint16_t* __restrict__ wb,
                                                                float* __restrict__ partials,
                                                                float* __restrict__ loss_part)
{
    static_assert(NT::fused && NT::W == 64, "the fused step is laid out for 64-wide hidden layers");
    constexpr int W = NT::W, L = NT::L, HS = W / 32;
    constexpr int T0 = NT::k(0) / 16 * (W / 16), TH = (W / 16) * (W / 16), TL = W / 16; // dW tiles per layer
    __shared__ __align__(16) uint16_t sWF[NT::frag_elems];
    __shared__ __align__(16) uint16_t sWB[NT::bfrag_elems];
    __shared__ __align__(16) uint16_t sH[2][kIn * kTrainLd];
    __shared__ __align__(16) uint16_t sZ[2][W * kTrainLd];
    __shared__ float sLoss[kTrainWaves];
    // the forward fragments are staged first; the gradient chain's are
    // loaded into registers behind the first pass's forward and stored
    // into LDS before its first barrier (the chain reads them only after it)
    constexpr int kWbVec = NT::bfrag_elems / 8, kWbPer = (kWbVec + 64 * kTrainWaves - 1) / (64 * kTrainWaves);
    static_assert(NT::bfrag_elems % 8 == 0, "16-byte pieces");
    stage16(sWF, wf, NT::frag_elems);
    __syncthreads();
    static_assert(kWbPer == 4, "four 16-byte pieces per thread");
    const uint4* wb4 = reinterpret_cast<const uint4*>(wb);
    constexpr int kStride = 64 * kTrainWaves;
    const int t0 = int(threadIdx.x);
    // in bounds: a piece past the end loads the last one (it is not stored)
    const uint4 wb0 = wb4[t0], wb1 = wb4[t0 + kStride], wb2 = wb4[t0 + 2 * kStride],
                wb3 = wb4[t0 + 3 * kStride < kWbVec ? t0 + 3 * kStride : kWbVec - 1];
    static_assert(3 * 64 * kTrainWaves <= kWbVec, "the first three pieces of every thread are in bounds");
    bool wb_pending = true;
    int off = 0, n = 0;
    span_rows(span, off, n);
    raw += size_t(off) * kRawStride;
    target += size_t(off) * 3;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
    const float inv_count = n > 0 ? 1.0f / (3.0f * float(n)) : 0.0f;
    const EncLane el = enc_lane(g);
    f4v a0[tiles_per_wave(T0)], ah[L - 1][tiles_per_wave(TH)], aL[tiles_per_wave(TL)];
#pragma unroll
    for (int i = 0; i < tiles_per_wave(T0); ++i) a0[i] = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int l = 0; l < L - 1; ++l)
#pragma unroll
        for (int i = 0; i < tiles_per_wave(TH); ++i) ah[l][i] = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int i = 0; i < tiles_per_wave(TL); ++i) aL[i] = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
    float loss = 0.0f;
    int buf = 0;
    const int passes = (n + kTrainRows - 1) / kTrainRows;
    for (int ps = blockIdx.x; ps < passes; ps += gridDim.x) {
        const int row = ps * kTrainRows + 16 * wave + c16;
        const bool ok = row < n;
        // forward: h[0] is layer 0's input, h[l] layer l's (l = 1..L)
        s8v h0[4][1], h[L + 1][4][1];
        enc_operands(raw + size_t(ok ? row : 0) * kRawStride, ok, g, el, h0[0][0], h0[1][0], h0[2][0]);
        {
            f4v acc[W / 16][1];
            qlayer<NT::kss(0), W / 16, 1>(sWF + NT::frag_off(0) * 512, h0, acc, lane);
            qactivate<W / 16, 1>(acc, h[1]);
        }
#pragma unroll
        for (int l = 1; l < L; ++l) {
            f4v acc[W / 16][1];
            qlayer<HS, W / 16, 1>(sWF + NT::frag_off(l) * 512, h[l], acc, lane);
            qactivate<W / 16, 1>(acc, h[l + 1]);
        }
        f4v y[1][1];
        qlayer<HS, 1, 1>(sWF + NT::frag_off(L) * 512, h[L], y, lane);
        // dz_L: dL/dy of channels 0..2 (lanes of group 0, slots 0..2), as the
        // B operand of the first gradient-chain product (features 0..15 in
        // slots 0..3 of k-step 0, the padding to 32 zero)
        float gy[3] = { 0.0f, 0.0f, 0.0f };
        if (g == 0 && ok) {
#pragma unroll
            for (int j = 0; j < 3; ++j) gy[j] = loss_grad(y[0][0][j], target[size_t(row) * 3 + j], inv_count, loss);
        }
        s8v dz[4][1];
        {
            const u4v d{ pack_bf16(gy[0], gy[1]), pack_bf16(gy[2], 0.0f), 0u, 0u };
            dz[0][0] = __builtin_bit_cast(s8v, d);
        }
        if (wb_pending) {
            uint4* s4 = reinterpret_cast<uint4*>(sWB);
            s4[t0] = wb0;
            s4[t0 + kStride] = wb1;
            s4[t0 + 2 * kStride] = wb2;
            if (t0 + 3 * kStride < kWbVec) s4[t0 + 3 * kStride] = wb3;
            wb_pending = false;
        }
        store_features<HS, W>(sH[buf], h[L], wave, lane);
        store_features<1, kOutCols>(sZ[buf], dz, wave, lane);
        __syncthreads();
        dw_pass<W / 16, 1>(sH[buf], sZ[buf], aL, wave, lane);
        buf ^= 1;
        // the chain: dz_{l-1} from dz_l, then dW_{l-1}
#pragma unroll
        for (int l = L; l >= 1; --l) {
            f4v acc[W / 16][1];
            static_assert(NT::bkss(L) == 1 && NT::bkss(1) == HS, "one k-step from the output layer's 16 columns");
            if (l == L)
                qlayer<1, W / 16, 1>(sWB + NT::bfrag_off(L) * 512, dz, acc, lane);
            else
                qlayer<HS, W / 16, 1>(sWB + NT::bfrag_off(l) * 512, dz, acc, lane);
#pragma unroll
            for (int ks = 0; ks < HS; ++ks) {
                const f4v lo = acc[2 * ks][0], hi = acc[2 * ks + 1][0];
                const u4v m = __builtin_bit_cast(u4v, h[l][ks][0]);
                const u4v d{ mask_positive(pack_bf16(lo[0], lo[1]), m[0]), mask_positive(pack_bf16(lo[2], lo[3]), m[1]),
                             mask_positive(pack_bf16(hi[0], hi[1]), m[2]), mask_positive(pack_bf16(hi[2], hi[3]), m[3]) };
                dz[ks][0] = __builtin_bit_cast(s8v, d);
            }
            if (l == 1)
                store_features<NT::kss(0), kIn>(sH[buf], h0, wave, lane);
            else
                store_features<HS, W>(sH[buf], h[l - 1], wave, lane);
            store_features<HS, W>(sZ[buf], dz, wave, lane);
            __syncthreads();
            if (l == 1)
                dw_pass<NT::k(0) / 16, W / 16>(sH[buf], sZ[buf], a0, wave, lane);
            else
                dw_pass<W / 16, W / 16>(sH[buf], sZ[buf], ah[l - 2], wave, lane);
            buf ^= 1;
        }
    }
    // this block's partial sums (blocks beyond the passes leave zeros)
    float* out = partials + size_t(blockIdx.x) * NT::params;
    dw_store<NT::k(0) / 16, W / 16>(out + NT::off(0), W, a0, wave, lane);
#pragma unroll
    for (int l = 1; l < L; ++l) dw_store<W / 16, W / 16>(out + NT::off(l), W, ah[l - 1], wave, lane);
    dw_store<W / 16, 1>(out + NT::off(L), kOutCols, aL, wave, lane);
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
    if (lane == 0) sLoss[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int w = 0; w < kTrainWaves; ++w) t += sLoss[w];
        loss_part[blockIdx.x] = t;
    }
}

} // namespace bb

// ================================================================== C ABI

struct BbNrc
{
    int device = 0;
    int width = 64;
    int hidden = 4;
    int params = 0;
    float lr = 1e-3f;
    int64_t step = 0;
    float* w = nullptr; // fp32 master
    float* m = nullptr;
    float* v = nullptr;
    float* g = nullptr;
    float* partials = nullptr; // [row range][params]
    float* loss = nullptr;
    float* loss_part = nullptr; // [block] loss partial sums of the last step
    int* status = nullptr; // [0]: diverged (non-finite loss, weight or prediction), sticky until new weights
    uint16_t* WT = nullptr;
    uint16_t* W = nullptr;
    uint16_t* WF = nullptr;   // weight fragments of the query kernel (and the fused forward)
    uint16_t* WB = nullptr;   // weight fragments of the fused gradient chain
    uint16_t* acts = nullptr; // saved layer inputs of the last training batch
    uint16_t* dzs = nullptr;  // saved layer-output gradients of the last training batch
    float* y = nullptr;
    size_t acts_rows = 0;
    bool diverged = false; // host copy of status[0] as of the last synchronizing call
};

namespace
{

bool nrc_ok(hipError_t e, const char* what)
{
    if (e == hipSuccess) return true;
    bb::set_error("%s: %s", what, hipGetErrorString(e));
    return false;
}

// blocks of a forward or gradient-chain launch over at most cap rows
// bytes of the query fragments (both shapes) and of the gradient-chain fragments
constexpr size_t kFragBytes = size_t(bb::NetC::frag_elems > bb::NetL::frag_elems ? bb::NetC::frag_elems : bb::NetL::frag_elems) * 2;
constexpr size_t kBFragBytes = size_t(bb::NetC::bfrag_elems) * 2;

int nrc_grid(int cap)
{
    const int waves = (cap + bb::kRows - 1) / bb::kRows;
    return std::max(1, std::min((waves + 3) / 4, 512));
}
static_assert(bb::kNrcBlock / 64 == 4, "nrc_grid assumes four waves per block");

int params_of(int width, int hidden)
{
    if (width == bb::NetC::W && hidden == bb::NetC::L) return bb::NetC::params;
    if (width == bb::NetL::W && hidden == bb::NetL::L) return bb::NetL::params;
    return -1;
}

// the Adam step (apply), or only the bf16 copies (reduce = 0: no gradient);
// nloss per-block loss sums, fused_grid as k_reduce_adam takes it
template <class NT>
bool nrc_reduce_t(BbNrc* x, hipStream_t st, bb::RowSpan span, bool reduce, int fused_grid, int nloss, bool apply,
                  float c1, float c2)
{
    hipLaunchKernelGGL(bb::k_reduce_adam<NT>, dim3((NT::params + 63) / 64), dim3(256), 0, st, x->w, x->m, x->v, x->g,
                       x->partials, span, reduce ? 1 : 0, fused_grid, x->lr, c1, c2, apply ? 1 : 0, x->WT, x->W, x->WF,
                       x->WB, x->loss_part, nloss, x->loss, x->status);
    return nrc_ok(hipGetLastError(), "k_reduce_adam");
}

bool nrc_reduce(BbNrc* x, hipStream_t st, bb::RowSpan span, bool reduce, int fused_grid, int nloss, bool apply,
                float c1, float c2)
{
    return x->width == bb::NetC::W ? nrc_reduce_t<bb::NetC>(x, st, span, reduce, fused_grid, nloss, apply, c1, c2)
                                   : nrc_reduce_t<bb::NetL>(x, st, span, reduce, fused_grid, nloss, apply, c1, c2);
}

// refreshes the bf16 copies from the fp32 master weights
bool nrc_refresh(BbNrc* x, hipStream_t st)
{
    return nrc_reduce(x, st, bb::RowSpan{ nullptr, 0, 0, 0 }, false, 0, 0, false, 1.0f, 1.0f);
}

// blocks of the query kernel that are resident at once on the device (one
// round): each block stages the weight fragments into LDS once, so a grid of
// one round, with the waves looping over the rows, stages them least (the
// row count is on the device; a grid sized for the capacity staged them up
// to 8 times per CU)
template <class NT>
int query_round()
{
    static const int blocks = [] { // computed once (thread-safe), on the first caller's device
        int dev = 0, cus = 0, per_cu = 0;
        bool failed = hipGetDevice(&dev) != hipSuccess;
        if (failed || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            failed = true;
            cus = 256;
        }
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bb::k_mlp_query<NT>, NT::qblock, 0) != hipSuccess) {
            failed = true;
            per_cu = 1;
        }
        if (failed) (void)hipGetLastError(); // the fallbacks above stand in; the launch reports its own error
        return std::max(cus, 1) * std::max(per_cu, 1);
    }();
    return blocks;
}

template <class NT>
void launch_query(BbNrc* x, const float* inputs_dev, bb::RowSpan span, float* out_dev, hipStream_t st)
{
    const long long per_block = (long long)(NT::qblock / 64) * 16 * NT::qtiles;
    const int grid =
        int(std::max<long long>(1, std::min<long long>((span.cap + per_block - 1) / per_block, query_round<NT>())));
    hipLaunchKernelGGL(bb::k_mlp_query<NT>, dim3(grid), dim3(NT::qblock), 0, st, inputs_dev, span, x->WF, out_dev);
}

int query_span(BbNrc* x, const float* inputs_dev, bb::RowSpan span, float* out_dev, hipStream_t st)
{
    if (x->width == bb::NetC::W)
        launch_query<bb::NetC>(x, inputs_dev, span, out_dev, st);
    else
        launch_query<bb::NetL>(x, inputs_dev, span, out_dev, st);
    return nrc_ok(hipGetLastError(), "k_mlp_query") ? BBGPU_OK : BBGPU_ERR_DEVICE;
}

// the saved activations, gradients and outputs of a step on up to rows rows
template <class NT>
bool ensure_train_buffers(BbNrc* x, int rows)
{
    if (x->acts_rows >= size_t(rows)) return true;
    const size_t chunks = (size_t(rows) + bb::kRows - 1) / bb::kRows;
    for (void* p : { (void*)x->acts, (void*)x->dzs, (void*)x->y })
        if (p) (void)hipFree(p);
    x->acts = nullptr;
    x->dzs = nullptr;
    x->y = nullptr;
    x->acts_rows = 0;
    if (!nrc_ok(hipMalloc(&x->acts, chunks * size_t(NT::chunk_acts) * 2), "hipMalloc") ||
        !nrc_ok(hipMalloc(&x->dzs, chunks * size_t(NT::chunk_dz) * 2), "hipMalloc") ||
        !nrc_ok(hipMalloc(&x->y, chunks * bb::kRows * 12), "hipMalloc"))
        return false;
    x->acts_rows = chunks * bb::kRows;
    return true;
}

// forward, gradient chain, dW partial sums and the reduction (with the Adam
// step when apply) of one step over the rows span selects; at most max_len
// rows (which sizes the launches).  The 64-wide network takes one fused
// kernel; the wide one the forward, gradient-chain and dW kernels, with its
// activations and gradients saved in between.
template <class NT>
int train_step(BbNrc* x, const float* inputs_dev, const float* targets_dev, bb::RowSpan span, int max_len, bool apply,
               hipStream_t st)
{
    int fused_grid = 0, nloss = 0;
    if constexpr (NT::fused) {
        fused_grid = bb::train_grid(max_len);
        nloss = fused_grid;
        hipLaunchKernelGGL(bb::k_mlp_train<NT>, dim3(fused_grid), dim3(64 * bb::kTrainWaves), 0, st, inputs_dev,
                           targets_dev, span, x->WF, x->WB, x->partials, x->loss_part);
        if (!nrc_ok(hipGetLastError(), "k_mlp_train")) return BBGPU_ERR_DEVICE;
    } else {
        if (!ensure_train_buffers<NT>(x, max_len)) return BBGPU_ERR_NOMEM;
        const int chunks = (max_len + bb::kRows - 1) / bb::kRows;
        const bb::NrcWeights Wt{ x->WT, x->W };
        const int grid = nrc_grid(max_len);
        nloss = grid;
        hipLaunchKernelGGL((bb::k_mlp_forward<NT, true>), dim3(grid), dim3(bb::kNrcBlock), 0, st, inputs_dev, span, Wt,
                           x->y, x->acts);
        hipLaunchKernelGGL(bb::k_mlp_dz<NT>, dim3(grid), dim3(bb::kNrcBlock), 0, st, x->y, targets_dev, span, x->W,
                           x->acts, x->dzs, x->loss_part);
        hipLaunchKernelGGL(bb::k_mlp_dw<NT>, dim3(NT::groups * bb::dw_splits(chunks, NT::groups)), dim3(64), 0, st,
                           x->acts, x->dzs, span, x->partials);
        if (!nrc_ok(hipGetLastError(), "k_mlp_dw")) return BBGPU_ERR_DEVICE;
    }
    // Adam bias corrections of the step about to be taken (every launched
    // step counts, also one that finds no rows and leaves the weights alone)
    float c1 = 1.0f, c2 = 1.0f;
    if (apply) {
        x->step++;
        c1 = 1.0f - std::pow(0.9f, float(x->step));
        c2 = 1.0f - std::pow(0.99f, float(x->step));
    }
    if (!nrc_reduce(x, st, span, true, fused_grid, nloss, apply, c1, c2)) return BBGPU_ERR_DEVICE;
    return BBGPU_OK;
}

int train_span(BbNrc* x, const float* inputs_dev, const float* targets_dev, bb::RowSpan span, int max_len, bool apply,
               hipStream_t st)
{
    return x->width == bb::NetC::W ? train_step<bb::NetC>(x, inputs_dev, targets_dev, span, max_len, apply, st)
                                   : train_step<bb::NetL>(x, inputs_dev, targets_dev, span, max_len, apply, st);
}

} // namespace

// Entry points for the cached renderer, whose row counts stay on the device
namespace bb
{

// predictions of min(*count, cap) rows
int nrc_query_counted(BbNrc* x, const float* inputs_dev, const int* count_dev, int cap, float* out_dev, hipStream_t st)
{
    return query_span(x, inputs_dev, RowSpan{ count_dev, cap, 0, 0 }, out_dev, st);
}

// Adam step number `step` of a pass over min(*count, cap) rows in batches of `batch`
int nrc_train_counted(BbNrc* x, const float* rows_dev, const float* targets_dev, const int* count_dev, int cap,
                      int batch, int step, hipStream_t st)
{
    const int max_len = std::max(1, std::min(batch, cap));
    return train_span(x, rows_dev, targets_dev, RowSpan{ count_dev, cap, batch, step }, max_len, true, st);
}

// device flag: the network diverged (set by the step, or by whoever applies its predictions)
int* nrc_status(BbNrc* x) { return x->status; }
bool nrc_diverged(const BbNrc* x) { return x->diverged; }
void nrc_note_diverged(BbNrc* x, bool d) { x->diverged = d; }
float* nrc_loss(BbNrc* x) { return x->loss; }

} // namespace bb

extern "C" {

int bbgpu_nrc_param_count(const BbNrcDesc* desc)
{
    return desc ? params_of(desc->hidden_width, desc->hidden_layers) : bb::NetC::params;
}

int bbgpu_nrc_create(const BbNrcDesc* desc, BbNrc** out)
{
    if (!desc || !out) {
        bb::set_error("bbgpu_nrc_create: null argument");
        return BBGPU_ERR_INVALID;
    }
    const int P = params_of(desc->hidden_width, desc->hidden_layers);
    if (P < 0) {
        bb::set_error("bbgpu_nrc_create: %d hidden layers of width %d are not built (%d x %d and %d x %d are)",
                      desc->hidden_layers, desc->hidden_width, bb::NetC::L, bb::NetC::W, bb::NetL::L, bb::NetL::W);
        return BBGPU_ERR_UNSUPPORTED;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        bb::set_error("bbgpu_nrc_create: no HIP device");
        return BBGPU_ERR_DEVICE;
    }
    BbNrc* x = new BbNrc();
    if (!nrc_ok(hipGetDevice(&x->device), "hipGetDevice")) {
        delete x;
        return BBGPU_ERR_DEVICE;
    }
    x->width = desc->hidden_width;
    x->hidden = desc->hidden_layers;
    x->params = P;
    x->lr = desc->learning_rate > 0.0f ? desc->learning_rate : 1e-3f;
    const size_t PS = size_t(P);
    bool ok = nrc_ok(hipMalloc(&x->w, PS * 4), "hipMalloc") && nrc_ok(hipMalloc(&x->m, PS * 4), "hipMalloc") &&
              nrc_ok(hipMalloc(&x->v, PS * 4), "hipMalloc") && nrc_ok(hipMalloc(&x->g, PS * 4), "hipMalloc") &&
              nrc_ok(hipMalloc(&x->loss, 4), "hipMalloc") && nrc_ok(hipMalloc(&x->status, 16), "hipMalloc") &&
              nrc_ok(hipMalloc(&x->WT, PS * 2), "hipMalloc") && nrc_ok(hipMalloc(&x->W, PS * 2), "hipMalloc") &&
              nrc_ok(hipMalloc(&x->WF, kFragBytes), "hipMalloc") && nrc_ok(hipMalloc(&x->WB, kBFragBytes), "hipMalloc") &&
              nrc_ok(hipMalloc(&x->loss_part, size_t(bb::kMaxSplits) * 4), "hipMalloc") &&
              nrc_ok(hipMalloc(&x->partials, size_t(bb::kMaxSplits) * PS * 4), "hipMalloc");
    if (!ok) {
        bbgpu_nrc_destroy(x);
        return BBGPU_ERR_NOMEM;
    }
    // Xavier-uniform initial weights from PCG32 seeded with desc->seed
    const int L = x->hidden, Wd = x->width;
    std::vector<float> w(PS, 0.0f);
    uint64_t st = bb::pcg_seed(uint64_t(desc->seed));
    size_t off = 0;
    for (int l = 0; l <= L; ++l) {
        const int K = l == 0 ? bb::kIn : Wd, N = l == L ? bb::kOutCols : Wd;
        const int fan_in = l == 0 ? 74 : K, fan_out = l == L ? 3 : N;
        const float a = std::sqrt(6.0f / float(fan_in + fan_out));
        for (int k = 0; k < K; ++k)
            for (int n = 0; n < N; ++n) {
                const float u = bb::pcg_float(st);
                const bool used = (l != L || n < 3) && (l != 0 || k < 74);
                w[off + size_t(k * N + n)] = used ? (2.0f * u - 1.0f) * a : 0.0f;
            }
        off += size_t(K) * N;
    }
    ok = nrc_ok(hipMemcpy(x->w, w.data(), PS * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         nrc_ok(hipMemset(x->m, 0, PS * 4), "hipMemset") && nrc_ok(hipMemset(x->v, 0, PS * 4), "hipMemset") &&
         nrc_ok(hipMemset(x->g, 0, PS * 4), "hipMemset") && nrc_ok(hipMemset(x->status, 0, 16), "hipMemset") &&
         nrc_ok(hipMemset(x->loss, 0, 4), "hipMemset") && nrc_ok(hipMemset(x->loss_part, 0, bb::kMaxSplits * 4), "hipMemset") &&
         // the fragments' padding (features beyond a layer's) stays zero
         nrc_ok(hipMemset(x->WF, 0, kFragBytes), "hipMemset") && nrc_ok(hipMemset(x->WB, 0, kBFragBytes), "hipMemset") &&
         nrc_refresh(x, nullptr) &&
         nrc_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (!ok) {
        bbgpu_nrc_destroy(x);
        return BBGPU_ERR_DEVICE;
    }
    *out = x;
    return BBGPU_OK;
}

int bbgpu_nrc_destroy(BbNrc* x)
{
    if (!x) return BBGPU_OK;
    for (void* p : { (void*)x->w, (void*)x->m, (void*)x->v, (void*)x->g, (void*)x->partials, (void*)x->loss,
                     (void*)x->loss_part, (void*)x->status, (void*)x->WT, (void*)x->W, (void*)x->WF, (void*)x->WB,
                     (void*)x->acts, (void*)x->dzs, (void*)x->y })
        if (p) (void)hipFree(p);
    delete x;
    return BBGPU_OK;
}

int bbgpu_nrc_get_weights(BbNrc* x, float* out, int32_t count)
{
    if (!x || !out || count != x->params) {
        bb::set_error("bbgpu_nrc_get_weights: invalid argument (count must be %d)", x ? x->params : 0);
        return BBGPU_ERR_INVALID;
    }
    if (!nrc_ok(hipMemcpy(out, x->w, size_t(count) * 4, hipMemcpyDeviceToHost), "hipMemcpy")) return BBGPU_ERR_DEVICE;
    return BBGPU_OK;
}

int bbgpu_nrc_set_weights(BbNrc* x, const float* in, int32_t count)
{
    if (!x || !in || count != x->params) {
        bb::set_error("bbgpu_nrc_set_weights: invalid argument (count must be %d)", x ? x->params : 0);
        return BBGPU_ERR_INVALID;
    }
    // new weights clear the divergence flag (non-finite ones set it again)
    int flag = 0;
    if (!nrc_ok(hipMemcpy(x->w, in, size_t(count) * 4, hipMemcpyHostToDevice), "hipMemcpy") ||
        !nrc_ok(hipMemset(x->status, 0, 16), "hipMemset") || !nrc_refresh(x, nullptr) ||
        !nrc_ok(hipMemcpy(&flag, x->status, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy"))
        return BBGPU_ERR_DEVICE;
    x->diverged = flag != 0;
    return BBGPU_OK;
}

int bbgpu_nrc_diverged(BbNrc* x, int32_t* diverged_out)
{
    if (!x || !diverged_out) {
        bb::set_error("bbgpu_nrc_diverged: null argument");
        return BBGPU_ERR_INVALID;
    }
    int flag = 0;
    if (!nrc_ok(hipMemcpy(&flag, x->status, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy")) return BBGPU_ERR_DEVICE;
    x->diverged = flag != 0;
    *diverged_out = flag != 0 ? 1 : 0;
    return BBGPU_OK;
}

int bbgpu_nrc_query(BbNrc* x, const float* inputs_dev, int32_t n, float* out_dev, void* stream)
{
    if (!x || n < 0 || (n > 0 && (!inputs_dev || !out_dev))) {
        bb::set_error("bbgpu_nrc_query: invalid argument");
        return BBGPU_ERR_INVALID;
    }
    if (n == 0) return BBGPU_OK;
    return query_span(x, inputs_dev, bb::RowSpan{ nullptr, n, 0, 0 }, out_dev, (hipStream_t)stream);
}

int bbgpu_nrc_train(BbNrc* x, const float* inputs_dev, const float* targets_dev, int32_t n, int32_t apply,
                    float* loss_out, float* grad_out, void* stream)
{
    if (!x || n <= 0 || !inputs_dev || !targets_dev) {
        bb::set_error("bbgpu_nrc_train: invalid argument");
        return BBGPU_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    // the gradient (the row ranges' partial sums), and the step when apply
    const int r = train_span(x, inputs_dev, targets_dev, bb::RowSpan{ nullptr, n, 0, 0 }, n, apply != 0, st);
    if (r != BBGPU_OK) return r;
    if (loss_out && !nrc_ok(hipMemcpyAsync(loss_out, x->loss, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync"))
        return BBGPU_ERR_DEVICE;
    if (grad_out &&
        !nrc_ok(hipMemcpyAsync(grad_out, x->g, size_t(x->params) * 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync"))
        return BBGPU_ERR_DEVICE;
    if ((loss_out || grad_out) && !nrc_ok(hipStreamSynchronize(st), "hipStreamSynchronize")) return BBGPU_ERR_DEVICE;
    return BBGPU_OK;
}

} // extern "C"
