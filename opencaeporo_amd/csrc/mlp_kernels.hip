// The learned form's network: a small MLP (P, T, z -> ln K_i and a
// stability logit) trained and run on the GPU with BF16 MFMA products and
// FP32 accumulation.  It only supplies the flash's initial K-values and a
// stability prior (ocpg_flash_fim_guided); the FP64 stability test and split
// still decide every cell.
//
//   mlpGemmKernel   : C = A * B (+ C) with A, B read in FP32 through
//                     arbitrary strides, rounded to BF16 and multiplied on
//                     gfx950's v_mfma_f32_16x16x32_bf16 (FP32 accumulators);
//                     one wavefront per 16 x 16 tile of C, the K loop
//                     optionally split across gridDim.z with FP32 partial
//                     tiles summed in a fixed order (deterministic).
//   mlpForwardKernel: the whole inference (4 layers, bias, GELU) in one
//                     kernel: the weights as BF16 MFMA operand fragments in
//                     LDS, then in registers; activations stay in registers
//                     between layers (the products are taken transposed, so
//                     one layer's accumulator layout is the next one's
//                     operand layout up to a fixed permutation of K).
//   bias/GELU, GELU', column sums, loss and Adam: elementwise / reductions.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "../../include/ocpgpu.h"
#include "mlp_kernels.hpp"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ __bf16 toBf16(float f)
{
    uint32_t u = __float_as_uint(f);
    uint16_t h;
    if ((u & 0x7fffffffu) > 0x7f800000u) {
        h = (uint16_t)((u >> 16) | 0x40);
    } else {
        const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
        h = (uint16_t)(r >> 16);
    }
    return __builtin_bit_cast(__bf16, h);
}

// One 16 x 16 tile of C per wavefront.  MFMA 16x16x32 operand layout
// (lane l, e = 0..7): A[i = l % 16][k = 8 (l / 16) + e], B[k = 8 (l / 16) + e]
// [n = l % 16]; result (j = 0..3): C[m = 4 (l / 16) + j][n = l % 16].
__global__ void __launch_bounds__(64) mlpGemmKernel(int M, int N, int K, const float* __restrict__ A, int64_t sAm,
                                                    int64_t sAk, const float* __restrict__ B, int64_t sBk,
                                                    int64_t sBn, float* __restrict__ C, int64_t sCm, int64_t sCn,
                                                    int accumulate, int kChunk, float* __restrict__ partial)
{
    const int lane = threadIdx.x;
    const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
    const int kb = blockIdx.z * kChunk;
    const int ke = min(K, kb + kChunk);
    const int row = lane & 15;
    const int k8 = (lane >> 4) * 8;
    const int kq = (lane >> 4) * 4;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + k8 + e;
            const int m = m0 + row, n = n0 + row;
            const float av = (m < M && k < ke) ? A[(int64_t)m * sAm + (int64_t)k * sAk] : 0.f;
            const float bv = (n < N && k < ke) ? B[(int64_t)k * sBk + (int64_t)n * sBn] : 0.f;
            a[e] = toBf16(av);
            b[e] = toBf16(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    const int n = n0 + row;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int m = m0 + kq + j;
        if (m >= M || n >= N) continue;
        if (partial) {
            partial[((int64_t)blockIdx.z * M + m) * N + n] = acc[j];
        } else {
            float* c = &C[(int64_t)m * sCm + (int64_t)n * sCn];
            *c = accumulate ? *c + acc[j] : acc[j];
        }
    }
}

// C (+)= sum over z of partial[z] (fixed order).
__global__ void mlpSplitSumKernel(int M, int N, int Z, const float* __restrict__ partial, float* __restrict__ C,
                                  int64_t sCm, int64_t sCn, int accumulate)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)M * N) return;
    const int m = (int)(t / N), n = (int)(t % N);
    float s = 0.f;
    for (int z = 0; z < Z; ++z) s += partial[((int64_t)z * M + m) * N + n];
    float* c = &C[(int64_t)m * sCm + (int64_t)n * sCn];
    *c = accumulate ? *c + s : s;
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float geluGrad(float x)
{
    const float cdf = 0.5f * (1.f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// Z[m][n] += b[n]; H = GELU(Z) (H may be NULL: bias only).
__global__ void mlpBiasActKernel(int M, int N, float* __restrict__ Z, const float* __restrict__ b,
                                 float* __restrict__ H)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)M * N) return;
    const float z = Z[t] + b[t % N];
    Z[t] = z;
    if (H) H[t] = gelu(z);
}

// dH *= GELU'(Z) in place.
__global__ void mlpActBackKernel(int64_t n, const float* __restrict__ Z, float* __restrict__ dH)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dH[t] *= geluGrad(Z[t]);
}

// out[n] = sum_m X[m][n], one block per column, fixed reduction order.
__global__ void mlpColSumKernel(int M, int N, const float* __restrict__ X, float* __restrict__ out)
{
    __shared__ float s[256];
    const int n = blockIdx.x;
    float v = 0.f;
    for (int m = threadIdx.x; m < M; m += blockDim.x) v += X[(int64_t)m * N + n];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = s[0];
}

// Loss per sample: mask * mean_i (Y_i - lnK_i)^2 + wStab * BCE(logit, label);
// dY (B x (nK + 1)) is its gradient divided by the batch size.  loss[b] gets
// the per-sample value.
__global__ void mlpLossKernel(int B, int nK, int ldY, const float* __restrict__ Y, const float* __restrict__ lnK,
                              const float* __restrict__ label, const float* __restrict__ mask, float wStab,
                              float* __restrict__ dY, float* __restrict__ loss)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float* y = Y + (int64_t)b * ldY;
    float* d = dY + (int64_t)b * ldY;
    const float mk = mask[b];
    float l = 0.f;
    for (int i = 0; i < nK; ++i) {
        const float e = y[i] - lnK[(int64_t)b * nK + i];
        l += mk * e * e / nK;
        d[i] = mk * 2.f * e / nK / B;
    }
    const float z = y[nK];
    const float p = 1.f / (1.f + expf(-z));
    const float t = label[b];
    l += wStab * (fmaxf(z, 0.f) - z * t + log1pf(expf(-fabsf(z))));
    d[nK] = wStab * (p - t) / B;
    for (int i = nK + 1; i < ldY; ++i) d[i] = 0.f;
    loss[b] = l;
}

__global__ void mlpAdamKernel(int64_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                              float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gi = g[t];
    const float mi = b1 * m[t] + (1.f - b1) * gi;
    const float vi = b2 * v[t] + (1.f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    w[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// Fused inference.  Per 16-row batch tile a wavefront computes
// Z_l^T = W_l^T H_{l-1}^T (M = the layer's outputs, N = 16 batch rows), so
// its accumulators hold (lane l, tile t, j = 0..3) Z[16 t + 4 (l / 16) + j]
// [row l % 16].  The next product's B operand needs, per K step s and lane
// l, 8 consecutive K positions 8 (l / 16) + e: they are taken as the
// hidden features f(g, e, s) = 16 (2 s + e / 4) + 4 g + e % 4 (g = l / 16),
// i.e. the lane's own accumulators of tiles 2 s and 2 s + 1, and the weight
// fragments (the A operand) are laid out with the same permutation.  The
// first layer reads X in the plain layout, K = the inputs padded to 32.
constexpr int kHid = 64;
constexpr int kFwdWaves = 4;
// A-operand fragments (64 lanes x 8 BF16) per layer: layer 1: 4 tiles x 1
// step, layers 2 and 3: 4 x 2, layer 4: 1 x 2.
constexpr int kFrag1 = 4, kFrag2 = 8, kFrag3 = 8, kFrag4 = 2;
constexpr int kFrags = kFrag1 + kFrag2 + kFrag3 + kFrag4;

__device__ __forceinline__ int hidOfK(int g, int e, int s) { return 16 * (2 * s + e / 4) + 4 * g + (e & 3); }

struct MlpWeights {
    const float *W1, *b1, *W2, *b2, *W3, *b3, *W4, *b4;  // W_l: (in x out) row-major
};

// The network's input: rows of X, or (X NULL) the features of cells of a
// SoA cell state: P / 5000, T / 600 and the hydrocarbon mole fractions
// Ni_i / sum_j |Ni_j| (j < NC), the flash's own feed.
struct MlpInput {
    const float* X;
    int ldX;
    const double *P, *T, *Ni;
    int64_t ld;
    int NC;
};

__device__ __forceinline__ float mlpFeature(const MlpInput& in, int64_t row, int k, int nin)
{
    if (k >= nin) return 0.f;
    if (in.X) return in.X[row * in.ldX + k];
    if (k == 0) return (float)(in.P[row] / 5000.0);
    if (k == 1) return (float)(in.T[row] / 600.0);
    double Nh = 0.0;
    for (int i = 0; i < in.NC; ++i) Nh += fabs(in.Ni[i * in.ld + row]);
    return (float)(in.Ni[(k - 2) * in.ld + row] / Nh);
}

__global__ void __launch_bounds__(64 * kFwdWaves) mlpForwardKernel(int B, int nin, int nout, MlpInput in,
                                                                   MlpWeights w, float* __restrict__ Y, int ldY)
{
    __shared__ bf16x8 frag[kFrags][64];
    __shared__ float bias[3 * kHid + 16];
    // stage the fragments: fragment q, lane l, element e
    for (int t = threadIdx.x; t < kFrags * 64; t += blockDim.x) {
        const int q = t / 64, l = t % 64, g = l / 16, i = l % 16;
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float x = 0.f;
            if (q < kFrag1) {                                   // layer 1, tile q
                const int k = 8 * g + e, o = 16 * q + i;
                if (k < nin) x = w.W1[k * kHid + o];
            } else if (q < kFrag1 + kFrag2 + kFrag3) {          // layers 2, 3: tile, step
                const int r = q < kFrag1 + kFrag2 ? q - kFrag1 : q - kFrag1 - kFrag2;
                const float* W = q < kFrag1 + kFrag2 ? w.W2 : w.W3;
                const int tile = r / 2, s = r % 2, o = 16 * tile + i;
                x = W[hidOfK(g, e, s) * kHid + o];
            } else {                                            // layer 4: step
                const int s = q - (kFrag1 + kFrag2 + kFrag3);
                if (i < nout) x = w.W4[hidOfK(g, e, s) * nout + i];
            }
            v[e] = toBf16(x);
        }
        frag[q][l] = v;
    }
    for (int t = threadIdx.x; t < 3 * kHid + 16; t += blockDim.x) {
        float x = 0.f;
        if (t < kHid) x = w.b1[t];
        else if (t < 2 * kHid) x = w.b2[t - kHid];
        else if (t < 3 * kHid) x = w.b3[t - 2 * kHid];
        else if (t - 3 * kHid < nout) x = w.b4[t - 3 * kHid];
        bias[t] = x;
    }
    __syncthreads();
    const int lane = threadIdx.x % 64, g = lane / 16, r = lane % 16;
    bf16x8 A[kFrags];
#pragma unroll
    for (int q = 0; q < kFrags; ++q) A[q] = frag[q][lane];
    float bl[3][4][4], b4[4];
#pragma unroll
    for (int L = 0; L < 3; ++L)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) bl[L][t][j] = bias[L * kHid + 16 * t + 4 * g + j];
#pragma unroll
    for (int j = 0; j < 4; ++j) b4[j] = bias[3 * kHid + 4 * g + j];
    const int tiles = (B + 15) / 16;
    const int wave = blockIdx.x * kFwdWaves + threadIdx.x / 64, waves = gridDim.x * kFwdWaves;
    for (int tile = wave; tile < tiles; tile += waves) {
        const int row = tile * 16 + r;
        const bool ok = row < B;
        // layer 1: B operand = X^T, K = the inputs
        bf16x8 xb;
#pragma unroll
        for (int e = 0; e < 8; ++e) xb[e] = toBf16(ok ? mlpFeature(in, row, 8 * g + e, nin) : 0.f);
        floatx4 z[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            z[t] = floatx4{0.f, 0.f, 0.f, 0.f};
            z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[t], xb, z[t], 0, 0, 0);
        }
        // layers 2 and 3
#pragma unroll
        for (int L = 0; L < 2; ++L) {
            bf16x8 hb[2];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int t = 2 * s + e / 4, j = e & 3;
                    hb[s][e] = toBf16(gelu(z[t][j] + bl[L][t][j]));
                }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                z[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[kFrag1 + L * kFrag2 + 2 * t + s], hb[s], z[t],
                                                                   0, 0, 0);
            }
        }
        // layer 4
        bf16x8 hb[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int t = 2 * s + e / 4, j = e & 3;
                hb[s][e] = toBf16(gelu(z[t][j] + bl[2][t][j]));
            }
        floatx4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; ++s)
            y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[kFrag1 + kFrag2 + kFrag3 + s], hb[s], y, 0, 0, 0);
        if (ok) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int o = 4 * g + j;
                if (o < nout) Y[(int64_t)row * ldY + o] = y[j] + b4[j];
            }
        }
    }
}

inline int status(hipError_t e) { return e == hipSuccess ? OCPG_OK : OCPG_EHIP; }
inline unsigned blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

namespace ocpg {

static hipError_t launchMlpForward(hipStream_t s, int B, int nin, int nout, const MlpInput& in,
                                   const float* const wb[8], float* Y, int ldY)
{
    const int tiles = (B + 15) / 16;
    // enough workgroups to fill the chip a few times over; each wavefront
    // loops over batch tiles, so the fragment staging is amortised
    const int wg = std::min((tiles + kFwdWaves - 1) / kFwdWaves, 2048);
    MlpWeights w{wb[0], wb[1], wb[2], wb[3], wb[4], wb[5], wb[6], wb[7]};
    hipLaunchKernelGGL(mlpForwardKernel, dim3(wg), dim3(64 * kFwdWaves), 0, s, B, nin, nout, in, w, Y, ldY);
    return hipGetLastError();
}

hipError_t launchMlpForwardCells(hipStream_t s, int64_t nb, int64_t ld, int NC, const double* P, const double* T,
                                 const double* Ni, const float* const wb[8], float* Y, int ldY)
{
    if (nb <= 0) return hipSuccess;
    MlpInput in{nullptr, 0, P, T, Ni, ld, NC};
    return launchMlpForward(s, (int)nb, NC + 2, NC + 1, in, wb, Y, ldY);
}

}  // namespace ocpg

// Split-K partial tiles of ocpg_mlp_gemm_noalloc, one copy per GPU (each
// device holds its own instance of this array from the first launch of the
// library's code on it until the process ends).  The host only ever passes
// the address hipGetSymbolAddress returns for the current device; the name
// itself is a host-side stand-in, not a device address.
static __device__ __align__(256) float scratch[OCPG_MLP_SCRATCH_FLOATS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex gScratchMu;
float* gScratchPtr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device.
int scratchOfCurrentDevice(float** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return OCPG_ENODEV;
    if (dev >= kMaxDevices) return OCPG_EUNSUP;
    std::lock_guard<std::mutex> g(gScratchMu);
    if (!gScratchPtr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return OCPG_EHIP;
        gScratchPtr[dev] = static_cast<float*>(p);
    }
    *out = gScratchPtr[dev];
    return OCPG_OK;
}

// Grid of mlpGemmKernel: the K chunk per z-slice and the number of slices.
void gemmSplit(int K, int splitK, int* kChunk, int* Z)
{
    *kChunk = ((K + splitK - 1) / splitK + 31) / 32 * 32;
    const int z = *kChunk > 0 ? (K + *kChunk - 1) / *kChunk : 0;
    *Z = z > 0 ? z : 1;
}

}  // namespace

extern "C" {

int ocpg_mlp_gemm(void* stream, int M, int N, int K, const float* A, int64_t sAm, int64_t sAk, const float* B,
                  int64_t sBk, int64_t sBn, float* C, int64_t sCm, int64_t sCn, int accumulate, int splitK,
                  float* partial)
{
    if (M <= 0 || N <= 0 || K < 0 || !A || !B || !C || splitK < 1 || (splitK > 1 && !partial)) return OCPG_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int kChunk, Z;
    gemmSplit(K, splitK, &kChunk, &Z);
    dim3 grid((M + 15) / 16, (N + 15) / 16, Z);
    hipLaunchKernelGGL(mlpGemmKernel, grid, dim3(64), 0, s, M, N, K, A, sAm, sAk, B, sBk, sBn, C, sCm, sCn,
                       accumulate, kChunk, Z > 1 ? partial : nullptr);
    if (Z > 1)
        hipLaunchKernelGGL(mlpSplitSumKernel, dim3(blocks((int64_t)M * N, 256)), dim3(256), 0, s, M, N, Z, partial,
                           C, sCm, sCn, accumulate);
    return status(hipGetLastError());
}

// ocpg_mlp_gemm with the split-K partial tiles in the library's static
// per-device buffer: allocates nothing.  The partial tiles need
// Z * M * N <= OCPG_MLP_SCRATCH_FLOATS floats (Z = the K slices actually
// used); larger requests are refused with OCPG_ETOOBIG before any launch.
int ocpg_mlp_gemm_noalloc(void* stream, int M, int N, int K, const float* A, int64_t sAm, int64_t sAk,
                          const float* B, int64_t sBk, int64_t sBn, float* C, int64_t sCm, int64_t sCn,
                          int accumulate, int splitK)
{
    if (M <= 0 || N <= 0 || K < 0 || !A || !B || !C || splitK < 1) return OCPG_EINVAL;
    int kChunk, Z;
    gemmSplit(K, splitK, &kChunk, &Z);
    if (Z > 1 && (int64_t)Z * M * N > (int64_t)OCPG_MLP_SCRATCH_FLOATS) return OCPG_ETOOBIG;
    float* part = nullptr;
    if (Z > 1) {
        const int rc = scratchOfCurrentDevice(&part);
        if (rc != OCPG_OK) return rc;
    }
    return ocpg_mlp_gemm(stream, M, N, K, A, sAm, sAk, B, sBk, sBn, C, sCm, sCn, accumulate, splitK, part);
}

// The network's inference in one kernel (see mlpForwardKernel): hidden
// width 64, nin <= 32 inputs, nout <= 16 outputs; the same BF16 roundings as
// the layer-by-layer path (operands rounded to BF16, FP32 accumulation, bias
// and GELU in FP32).
int ocpg_mlp_forward(void* stream, int B, int nin, int nout, const float* X, int ldX, const float* W1,
                     const float* b1, const float* W2, const float* b2, const float* W3, const float* b3,
                     const float* W4, const float* b4, float* Y, int ldY)
{
    if (B <= 0 || nin < 1 || nin > 32 || nout < 1 || nout > 16 || ldX < nin || ldY < nout || !X || !Y || !W1 ||
        !b1 || !W2 || !b2 || !W3 || !b3 || !W4 || !b4)
        return OCPG_EINVAL;
    MlpInput in{X, ldX, nullptr, nullptr, nullptr, 0, 0};
    const float* wb[8] = {W1, b1, W2, b2, W3, b3, W4, b4};
    return status(ocpg::launchMlpForward(static_cast<hipStream_t>(stream), B, nin, nout, in, wb, Y, ldY));
}

int ocpg_mlp_bias_act(void* stream, int M, int N, float* Z, const float* b, float* H)
{
    if (M <= 0 || N <= 0 || !Z || !b) return OCPG_EINVAL;
    hipLaunchKernelGGL(mlpBiasActKernel, dim3(blocks((int64_t)M * N, 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), M, N, Z, b, H);
    return status(hipGetLastError());
}

int ocpg_mlp_act_backward(void* stream, int64_t n, const float* Z, float* dH)
{
    if (n <= 0 || !Z || !dH) return OCPG_EINVAL;
    hipLaunchKernelGGL(mlpActBackKernel, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), n,
                       Z, dH);
    return status(hipGetLastError());
}

int ocpg_mlp_colsum(void* stream, int M, int N, const float* X, float* out)
{
    if (M <= 0 || N <= 0 || !X || !out) return OCPG_EINVAL;
    hipLaunchKernelGGL(mlpColSumKernel, dim3(N), dim3(256), 0, static_cast<hipStream_t>(stream), M, N, X, out);
    return status(hipGetLastError());
}

int ocpg_mlp_loss(void* stream, int B, int nK, int ldY, const float* Y, const float* lnK, const float* label,
                  const float* mask, float wStab, float* dY, float* loss)
{
    if (B <= 0 || nK <= 0 || ldY < nK + 1 || !Y || !lnK || !label || !mask || !dY || !loss) return OCPG_EINVAL;
    hipLaunchKernelGGL(mlpLossKernel, dim3(blocks(B, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), B, nK,
                       ldY, Y, lnK, label, mask, wStab, dY, loss);
    return status(hipGetLastError());
}

int ocpg_mlp_adam(void* stream, int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1,
                  float b2, float eps, int step)
{
    if (n <= 0 || !w || !g || !m || !v || step < 1) return OCPG_EINVAL;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(mlpAdamKernel, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), n, w,
                       g, m, v, lr, b1, b2, eps, c1, c2);
    return status(hipGetLastError());
}

}  // extern "C"
