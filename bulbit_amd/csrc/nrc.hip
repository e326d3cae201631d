// nrc.hip — the radiance cache: a small MLP evaluated and trained with bf16
// MFMA on gfx950 (fp32 accumulation, fp32 master weights, Adam).
//
// Network (no biases): the 74-wide input encoding padded to kIn = 80,
// kHidden hidden layers of width kWidth with ReLU, and a linear output of 3
// (stored as 16 columns, the last 13 held at zero):
//   h0 = enc(x);  h_{l+1} = relu(h_l W_l) for l < kHidden;  y = h_kHidden W_kHidden
// Input encoding of a raw row (16 floats: position in [0,1]^3, direction,
// normal, roughness, albedo): sin/cos(2^f pi p_a) frequency-major for
// f = 0..11 and axis a (72 values, the first 64 kept), then direction (3),
// normal (3), roughness (1), albedo (3), zeros to 80.
//
// Matrix products use v_mfma_f32_16x16x16_bf16: a 16x16 tile of A (rows x k)
// comes from lane l as A[l % 16][4 (l / 16) + j], j = 0..3; B (k x n) as
// B[4 (l / 16) + j][l % 16]; the f32 result D[4 (l / 16) + j][l % 16].
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

constexpr int kIn = 80;      // encoded input width (74 used)
constexpr int kWidth = 64;   // hidden width
constexpr int kHidden = 4;   // hidden layers
constexpr int kOutCols = 16; // output columns (3 used)
constexpr int kRawStride = 16;
constexpr int kRows = 64;    // rows per wave per chunk (4 row tiles of 16)
constexpr int kNrcBlock = 256;

// weight matrices: W_0 [kIn][kWidth], W_1..W_{kHidden-1} [kWidth][kWidth],
// W_kHidden [kWidth][kOutCols]; fp32 master copies k-major ([k][n])
constexpr int layer_k(int l) { return l == 0 ? kIn : kWidth; }
constexpr int layer_n(int l) { return l == kHidden ? kOutCols : kWidth; }
constexpr int layer_off(int l) { return l == 0 ? 0 : layer_off(l - 1) + layer_k(l - 1) * layer_n(l - 1); }
constexpr int kParams = layer_off(kHidden + 1);

// Layer inputs saved by the training forward pass, per chunk of kRows rows,
// feature-major ([feature][row in chunk]): h_0 (kIn features), then
// h_1..h_kHidden (kWidth each).  This is the operand layout the backward
// MFMAs read (four consecutive rows of one feature in 8 bytes).
constexpr int act_off(int l) { return l == 0 ? 0 : kRows * kIn + (l - 1) * kRows * kWidth; }
constexpr int kChunkActs = act_off(kHidden + 1);

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

// --------------------------------------------------------------- encoding

BB_D void encode_row(const float* r, uint16_t* out)
{
    int k = 0;
    for (int f = 0; f < 12 && k < 64; ++f) {
        const float s = float(1 << f);
        for (int a = 0; a < 3 && k < 64; ++a) {
            out[k++] = to_bf16(sinpif(s * r[a]));
            out[k++] = to_bf16(cospif(s * r[a]));
        }
    }
    for (int i = 3; i < 13; ++i) out[k++] = to_bf16(r[i]);
    while (k < kIn) out[k++] = 0;
}

// ------------------------------------------------------------ MFMA pieces

BB_D s4 load4(const uint16_t* p) { return *reinterpret_cast<const s4*>(p); }

BB_D s4 gather4(const uint16_t* p, int stride)
{
    s4 v;
    v.x = short(p[0]);
    v.y = short(p[stride]);
    v.z = short(p[2 * stride]);
    v.w = short(p[3 * stride]);
    return v;
}

BB_D f4v mfma(s4 a, s4 b, f4v c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// one layer for a wave's kRows rows: act [kRows][ld_in] (bf16, LDS) times
// W_l (WT: [n][k] bf16) -> acc[row tile][col tile]
template <int K, int N>
BB_D void layer_forward(const uint16_t* act, int ld_in, const uint16_t* WT, f4v (&acc)[4][N / 16], int lane)
{
    const int r16 = lane & 15, q = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < N / 16; ++ct) acc[rt][ct] = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int kk = 0; kk < K / 16; ++kk) {
        s4 b[N / 16];
#pragma unroll
        for (int ct = 0; ct < N / 16; ++ct) b[ct] = load4(WT + (ct * 16 + r16) * K + kk * 16 + q * 4);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const s4 a = load4(act + (rt * 16 + r16) * ld_in + kk * 16 + q * 4);
#pragma unroll
            for (int ct = 0; ct < N / 16; ++ct) acc[rt][ct] = mfma(a, b[ct], acc[rt][ct]);
        }
    }
}

// ---------------------------------------------------------------- forward

struct NrcWeights
{
    const uint16_t* WT; // bf16 [n][k] per layer (at layer_off)
    const uint16_t* W;  // bf16 [k][n] per layer
};

// y = f(enc(raw)) for n rows; with SAVE the bf16 input of every layer is
// stored to acts[l][row][0..K_l) (layer l's input h_l) for the backward pass
template <bool SAVE>
__global__ __launch_bounds__(kNrcBlock) void k_mlp_forward(const float* __restrict__ raw, int n, NrcWeights Wt,
                                                            float* __restrict__ out, uint16_t* __restrict__ acts)
{
    __shared__ __align__(16) uint16_t sW[kParams];
    __shared__ uint16_t sAct[kNrcBlock / 64][kRows * kIn];
    static_assert(kParams % 8 == 0, "weights staged 16 bytes at a time");
    for (int i = threadIdx.x; i < kParams / 8; i += blockDim.x)
        reinterpret_cast<uint4*>(sW)[i] = reinterpret_cast<const uint4*>(Wt.WT)[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* act = sAct[wave];
    const int r16 = lane & 15, q = lane >> 4;
    const int chunks = (n + kRows - 1) / kRows;
    for (int ch = blockIdx.x * (kNrcBlock / 64) + wave; ch < chunks; ch += gridDim.x * (kNrcBlock / 64)) {
        const int row0 = ch * kRows;
        {
            // each lane encodes its row straight into the wave's LDS rows
            const int row = row0 + lane;
            uint16_t* e = act + lane * kIn;
            if (row < n)
                encode_row(raw + size_t(row) * kRawStride, e);
            else
                for (int i = 0; i < kIn; ++i) e[i] = 0;
            if (SAVE) // every row of the chunk, zeros beyond n
                for (int i = 0; i < kIn; ++i) acts[size_t(ch) * kChunkActs + act_off(0) + i * kRows + lane] = e[i];
        }
        // hidden layers: layer 0 (kIn inputs), then the kWidth-input ones in a
        // rolled loop (one copy of the code: small kernels stay in the
        // instruction cache when every wave runs a single chunk)
#pragma unroll 1
        for (int l = 0; l < kHidden; ++l) {
            f4v acc[4][kWidth / 16];
            if (l == 0)
                layer_forward<kIn, kWidth>(act, kIn, sW + layer_off(0), acc, lane);
            else
                layer_forward<kWidth, kWidth>(act, kIn, sW + layer_off(l), acc, lane);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                for (int ct = 0; ct < kWidth / 16; ++ct) {
                    s4 hv;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int r = rt * 16 + q * 4 + j, c = ct * 16 + r16;
                        const uint16_t h = to_bf16(fmaxf(acc[rt][ct][j], 0.0f));
                        act[r * kIn + c] = h;
                        hv[j] = short(h);
                    }
                    // rows rt*16 + 4q .. +3 of feature ct*16 + r16 of h_{l+1}
                    if (SAVE)
                        *reinterpret_cast<s4*>(acts + size_t(ch) * kChunkActs + act_off(l + 1) +
                                               (ct * 16 + r16) * kRows + rt * 16 + q * 4) = hv;
                }
        }
        // output layer
        f4v acc[4][1];
        layer_forward<kWidth, kOutCols>(act, kIn, sW + layer_off(kHidden), acc, lane);
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

// One layer of the backward pass for a wave's chunk: dW_l += h_l^T dz (K x N)
// with dz in registers (dzb, the MFMA result layout) and row-major in dzl;
// with DH, dz becomes dh_l = dz W_l^T masked by h_l > 0 (kWidth columns).
template <int K, int N, bool DH>
BB_D void bwd_layer(int l, const uint16_t* hc, const NrcWeights& Wt, float* sG, uint16_t* dzl,
                    s4 (&dzb)[4][kWidth / 16], int lane)
{
    const int r16 = lane & 15, q = lane >> 4;
    const uint16_t* h = hc + act_off(l);
    float* g = sG + layer_off(l);
    // every global operand of the layer first, so that the wave (alone on
    // its SIMD) waits for memory once per layer: h^T tiles, the ReLU mask
    // (h_l itself) and W_l
    s4 a[K / 16][4];
#pragma unroll
    for (int mt = 0; mt < K / 16; ++mt)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) a[mt][kk] = load4(h + (mt * 16 + r16) * kRows + kk * 16 + q * 4);
    s4 hv[DH ? 4 : 1][kWidth / 16];
    s4 b[DH ? N / 16 : 1][kWidth / 16];
    if constexpr (DH) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < kWidth / 16; ++ct) hv[rt][ct] = load4(h + (ct * 16 + r16) * kRows + rt * 16 + q * 4);
#pragma unroll
        for (int kk = 0; kk < N / 16; ++kk)
#pragma unroll
            for (int ct = 0; ct < kWidth / 16; ++ct)
                b[kk][ct] = load4(Wt.W + layer_off(l) + (ct * 16 + r16) * N + kk * 16 + q * 4);
    }
    // dW_l += h_l^T dz: A = h^T (features x rows), B = dz (rows x N)
#pragma unroll
    for (int mt = 0; mt < K / 16; ++mt)
#pragma unroll
        for (int nt = 0; nt < N / 16; ++nt) {
            f4v acc = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) acc = mfma(a[mt][kk], dzb[kk][nt], acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) atomicAdd(&g[(mt * 16 + q * 4 + j) * N + nt * 16 + r16], acc[j]);
        }
    if constexpr (!DH) return;
    // dh_l = dz W_l^T: (rows x N) (N x kWidth)
    f4v acc[4][kWidth / 16];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < kWidth / 16; ++ct) acc[rt][ct] = f4v{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int kk = 0; kk < N / 16; ++kk) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const s4 av = load4(dzl + (rt * 16 + r16) * kWidth + kk * 16 + q * 4);
#pragma unroll
            for (int ct = 0; ct < kWidth / 16; ++ct) acc[rt][ct] = mfma(av, b[kk][ct], acc[rt][ct]);
        }
    }
    // mask by h_l > 0 and keep dz_l in both forms
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < kWidth / 16; ++ct) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint16_t hj = uint16_t(hv[rt][ct][j]);
                const bool on = hj != 0 && !(hj & 0x8000u);
                const uint16_t d = on ? to_bf16(acc[rt][ct][j]) : uint16_t(0);
                dzb[rt][ct][j] = short(d);
                dzl[(rt * 16 + q * 4 + j) * kWidth + ct * 16 + r16] = d;
            }
        }
}

// Gradient of the loss for n rows whose layer inputs the forward pass saved
// (feature-major per chunk): dW_l = h_l^T dz_{l+1}, dh_l = dz_{l+1} W_l^T
// masked by h_l > 0.  Per chunk of 64 rows a wave keeps dz in registers in
// the MFMA result layout, which is exactly the B operand of dW = h^T dz (the
// lane's own tile), and a row-major bf16 copy in LDS for the A operand of
// dz W^T; h^T and the ReLU mask come from the saved layout with one 8-byte
// load each.  Each block sums its waves' dW in LDS and stores the sum to
// partials[block] (k_adam adds the blocks in order).
__global__ __launch_bounds__(kNrcBlock) void k_mlp_backward(const float* __restrict__ y,
                                                             const float* __restrict__ target, int n,
                                                             NrcWeights Wt, const uint16_t* __restrict__ acts,
                                                             float* __restrict__ partials, float* __restrict__ loss_sum)
{
    __shared__ float sG[kParams];
    __shared__ uint16_t sD[kNrcBlock / 64][kRows * kWidth]; // dz of the current layer output, row-major
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) sG[i] = 0.0f;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, q = lane >> 4;
    uint16_t* dzl = sD[wave];
    const float inv_count = 1.0f / (3.0f * float(n));
    float loss = 0.0f;
    const int chunks = (n + kRows - 1) / kRows;
    for (int ch = blockIdx.x * (kNrcBlock / 64) + wave; ch < chunks; ch += gridDim.x * (kNrcBlock / 64)) {
        const int row0 = ch * kRows;
        const uint16_t* hc = acts + size_t(ch) * kChunkActs;
        // dz of the output: dL/dy of rows rt*16 + 4q + j, column r16 (3 used)
        s4 dzb[4][kWidth / 16];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = rt * 16 + q * 4 + j, row = row0 + r;
                float g = 0.0f;
                if (r16 < 3 && row < n) g = loss_grad(y[size_t(row) * 3 + r16], target[size_t(row) * 3 + r16], inv_count, loss);
                const uint16_t gb = to_bf16(g);
                dzb[rt][0][j] = short(gb);
                dzl[r * kWidth + r16] = gb;
            }
        }
        // layers from the output down; the kWidth x kWidth ones in a rolled loop
        bwd_layer<kWidth, kOutCols, true>(kHidden, hc, Wt, sG, dzl, dzb, lane);
#pragma unroll 1
        for (int l = kHidden - 1; l >= 1; --l) bwd_layer<kWidth, kWidth, true>(l, hc, Wt, sG, dzl, dzb, lane);
        bwd_layer<kIn, kWidth, false>(0, hc, Wt, sG, dzl, dzb, lane);
    }
    // wave loss -> global
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (lane == 0) atomicAdd(loss_sum, loss);
    __syncthreads();
    float* out = partials + size_t(blockIdx.x) * kParams;
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) out[i] = sG[i];
}

// The gradient as the sum of the backward blocks' partial sums (in block
// order), kept in g; then Adam (beta1 0.9, beta2 0.99, eps 1e-8) on the fp32
// master weights when apply, and the bf16 copies in both layouts; the unused
// output columns stay zero
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, float* __restrict__ g,
                       const float* __restrict__ partials, int n_partials, float lr, float c1, float c2, int apply,
                       uint16_t* __restrict__ WT, uint16_t* __restrict__ W)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    int l = 0;
    while (l < kHidden && i >= layer_off(l + 1)) ++l;
    const int K = layer_k(l), N = layer_n(l);
    const int k = (i - layer_off(l)) / N, nn = (i - layer_off(l)) % N;
    float gi = 0.0f;
    for (int b = 0; b < n_partials; ++b) gi += partials[size_t(b) * kParams + i];
    if (n_partials > 0) g[i] = gi;
    float wi = w[i];
    if (apply && !(l == kHidden && nn >= 3)) {
        const float mi = 0.9f * m[i] + 0.1f * gi;
        const float vi = 0.99f * v[i] + 0.01f * gi * gi;
        m[i] = mi;
        v[i] = vi;
        wi -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
        w[i] = wi;
    }
    const uint16_t h = to_bf16(wi);
    W[i] = h;
    WT[layer_off(l) + nn * K + k] = h;
}

__global__ void k_fill0(float* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

} // namespace bb

// ================================================================== C ABI

struct BbNrc
{
    int device = 0;
    float lr = 1e-3f;
    int64_t step = 0;
    float* w = nullptr; // fp32 master
    float* m = nullptr;
    float* v = nullptr;
    float* g = nullptr;
    float* partials = nullptr; // [backward block][kParams]
    float* loss = nullptr;
    uint16_t* WT = nullptr;
    uint16_t* W = nullptr;
    uint16_t* acts = nullptr; // saved layer inputs of the last training batch
    float* y = nullptr;
    size_t acts_rows = 0;
};

namespace
{

bool nrc_ok(hipError_t e, const char* what)
{
    if (e == hipSuccess) return true;
    bb::set_error("%s: %s", what, hipGetErrorString(e));
    return false;
}

int nrc_grid(int n)
{
    const int waves = (n + bb::kRows - 1) / bb::kRows;
    return std::max(1, std::min((waves + 3) / 4, 512));
}
static_assert(bb::kNrcBlock / 64 == 4, "nrc_grid assumes four waves per block");

constexpr int kMaxBackwardBlocks = 512;

bool nrc_sync_bf16(BbNrc* x, hipStream_t st, bool apply, float c1, float c2, int n_partials = 0)
{
    hipLaunchKernelGGL(bb::k_adam, dim3((bb::kParams + 255) / 256), dim3(256), 0, st, x->w, x->m, x->v, x->g,
                       x->partials, n_partials, x->lr, c1, c2, apply ? 1 : 0, x->WT, x->W);
    return nrc_ok(hipGetLastError(), "k_adam");
}

} // namespace

extern "C" {

int bbgpu_nrc_param_count(void) { return bb::kParams; }

int bbgpu_nrc_create(const BbNrcDesc* desc, BbNrc** out)
{
    if (!desc || !out) {
        bb::set_error("bbgpu_nrc_create: null argument");
        return BBGPU_ERR_INVALID;
    }
    if (desc->hidden_width != bb::kWidth || desc->hidden_layers != bb::kHidden) {
        bb::set_error("bbgpu_nrc_create: only %d hidden layers of width %d are built", bb::kHidden, bb::kWidth);
        return BBGPU_ERR_UNSUPPORTED;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        bb::set_error("bbgpu_nrc_create: no HIP device");
        return BBGPU_ERR_DEVICE;
    }
    BbNrc* x = new BbNrc();
    if (hipGetDevice(&x->device) != hipSuccess) x->device = 0;
    x->lr = desc->learning_rate > 0.0f ? desc->learning_rate : 1e-3f;
    const size_t P = bb::kParams;
    bool ok = nrc_ok(hipMalloc(&x->w, P * 4), "hipMalloc") && nrc_ok(hipMalloc(&x->m, P * 4), "hipMalloc") &&
              nrc_ok(hipMalloc(&x->v, P * 4), "hipMalloc") && nrc_ok(hipMalloc(&x->g, P * 4), "hipMalloc") &&
              nrc_ok(hipMalloc(&x->loss, 4), "hipMalloc") && nrc_ok(hipMalloc(&x->WT, P * 2), "hipMalloc") &&
              nrc_ok(hipMalloc(&x->W, P * 2), "hipMalloc") &&
              nrc_ok(hipMalloc(&x->partials, size_t(kMaxBackwardBlocks) * P * 4), "hipMalloc");
    if (!ok) {
        bbgpu_nrc_destroy(x);
        return BBGPU_ERR_NOMEM;
    }
    // Xavier-uniform initial weights from PCG32 seeded with desc->seed
    std::vector<float> w(P, 0.0f);
    uint64_t st = bb::pcg_seed(uint64_t(desc->seed));
    for (int l = 0; l <= bb::kHidden; ++l) {
        const int K = bb::layer_k(l), N = bb::layer_n(l);
        const int fan_in = l == 0 ? 74 : K, fan_out = l == bb::kHidden ? 3 : N;
        const float a = std::sqrt(6.0f / float(fan_in + fan_out));
        for (int k = 0; k < K; ++k)
            for (int n = 0; n < N; ++n) {
                const float u = bb::pcg_float(st);
                const bool used = (l != bb::kHidden || n < 3) && (l != 0 || k < 74);
                w[size_t(bb::layer_off(l) + k * N + n)] = used ? (2.0f * u - 1.0f) * a : 0.0f;
            }
    }
    ok = nrc_ok(hipMemcpy(x->w, w.data(), P * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
         nrc_ok(hipMemset(x->m, 0, P * 4), "hipMemset") && nrc_ok(hipMemset(x->v, 0, P * 4), "hipMemset") &&
         nrc_ok(hipMemset(x->g, 0, P * 4), "hipMemset") && nrc_sync_bf16(x, nullptr, false, 1.0f, 1.0f) &&
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
                     (void*)x->WT, (void*)x->W, (void*)x->acts, (void*)x->y })
        if (p) (void)hipFree(p);
    delete x;
    return BBGPU_OK;
}

int bbgpu_nrc_get_weights(BbNrc* x, float* out, int32_t count)
{
    if (!x || !out || count != bb::kParams) {
        bb::set_error("bbgpu_nrc_get_weights: invalid argument (count must be %d)", bb::kParams);
        return BBGPU_ERR_INVALID;
    }
    if (!nrc_ok(hipMemcpy(out, x->w, size_t(count) * 4, hipMemcpyDeviceToHost), "hipMemcpy")) return BBGPU_ERR_DEVICE;
    return BBGPU_OK;
}

int bbgpu_nrc_set_weights(BbNrc* x, const float* in, int32_t count)
{
    if (!x || !in || count != bb::kParams) {
        bb::set_error("bbgpu_nrc_set_weights: invalid argument (count must be %d)", bb::kParams);
        return BBGPU_ERR_INVALID;
    }
    if (!nrc_ok(hipMemcpy(x->w, in, size_t(count) * 4, hipMemcpyHostToDevice), "hipMemcpy") ||
        !nrc_sync_bf16(x, nullptr, false, 1.0f, 1.0f) || !nrc_ok(hipDeviceSynchronize(), "hipDeviceSynchronize"))
        return BBGPU_ERR_DEVICE;
    return BBGPU_OK;
}

int bbgpu_nrc_query(BbNrc* x, const float* inputs_dev, int32_t n, float* out_dev, void* stream)
{
    if (!x || n < 0 || (n > 0 && (!inputs_dev || !out_dev))) {
        bb::set_error("bbgpu_nrc_query: invalid argument");
        return BBGPU_ERR_INVALID;
    }
    if (n == 0) return BBGPU_OK;
    hipStream_t st = (hipStream_t)stream;
    const bb::NrcWeights Wt{ x->WT, x->W };
    hipLaunchKernelGGL(bb::k_mlp_forward<false>, dim3(nrc_grid(n)), dim3(bb::kNrcBlock), 0, st, inputs_dev, n, Wt,
                       out_dev, nullptr);
    return nrc_ok(hipGetLastError(), "k_mlp_forward") ? BBGPU_OK : BBGPU_ERR_DEVICE;
}

int bbgpu_nrc_train(BbNrc* x, const float* inputs_dev, const float* targets_dev, int32_t n, int32_t apply,
                    float* loss_out, float* grad_out, void* stream)
{
    if (!x || n <= 0 || !inputs_dev || !targets_dev) {
        bb::set_error("bbgpu_nrc_train: invalid argument");
        return BBGPU_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    const size_t chunks = (size_t(n) + bb::kRows - 1) / bb::kRows;
    const size_t need = chunks * size_t(bb::kChunkActs);
    if (x->acts_rows < size_t(n)) {
        if (x->acts) (void)hipFree(x->acts);
        if (x->y) (void)hipFree(x->y);
        x->acts = nullptr;
        x->y = nullptr;
        x->acts_rows = 0;
        if (!nrc_ok(hipMalloc(&x->acts, need * 2), "hipMalloc") || !nrc_ok(hipMalloc(&x->y, size_t(n) * 12), "hipMalloc"))
            return BBGPU_ERR_NOMEM;
        x->acts_rows = size_t(n);
    }
    const bb::NrcWeights Wt{ x->WT, x->W };
    const int grid = nrc_grid(n);
    hipLaunchKernelGGL(bb::k_fill0, dim3(1), dim3(64), 0, st, x->loss, 1);
    hipLaunchKernelGGL(bb::k_mlp_forward<true>, dim3(grid), dim3(bb::kNrcBlock), 0, st, inputs_dev, n, Wt, x->y,
                       x->acts);
    hipLaunchKernelGGL(bb::k_mlp_backward, dim3(grid), dim3(bb::kNrcBlock), 0, st, x->y, targets_dev, n, Wt, x->acts,
                       x->partials, x->loss);
    if (!nrc_ok(hipGetLastError(), "k_mlp_backward")) return BBGPU_ERR_DEVICE;
    // the gradient (partial sums of the grid's blocks), and the step when apply
    float c1 = 1.0f, c2 = 1.0f;
    if (apply) {
        x->step++;
        c1 = 1.0f - std::pow(0.9f, float(x->step));
        c2 = 1.0f - std::pow(0.99f, float(x->step));
    }
    if (!nrc_sync_bf16(x, st, apply != 0, c1, c2, grid)) return BBGPU_ERR_DEVICE;
    if (loss_out && !nrc_ok(hipMemcpyAsync(loss_out, x->loss, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync"))
        return BBGPU_ERR_DEVICE;
    if (grad_out &&
        !nrc_ok(hipMemcpyAsync(grad_out, x->g, size_t(bb::kParams) * 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync"))
        return BBGPU_ERR_DEVICE;
    if ((loss_out || grad_out) && !nrc_ok(hipStreamSynchronize(st), "hipStreamSynchronize")) return BBGPU_ERR_DEVICE;
    return BBGPU_OK;
}

} // extern "C"
