// Learned far-field correction: a 12 -> 64 -> 64 -> 64 -> 9 MLP (smooth
// hard-SiLU activations, below) on BF16
// MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation) that predicts the
// residual between the exact force of a cell's bodies and the cell's
// multipole (monopole + quadrupole, softened as in grav.c:42-73) force at a
// point, and the residual's Jacobian there (symmetric: xx xy xz yy yz zz), so
// that the correction of a group member at offset dx from the group centre is
// R + J dx.
//
// One wavefront handles 64 samples (rows).  A layer is a 64 x out x in GEMM
// cut into 16x16 output tiles, each the sum of in/32 MFMAs.  Operand layout
// of the 16x16x32 instruction (wave64): lane l holds A[l%16][8*(l/16)+e] and
// B[8*(l/16)+e][l%16] for e = 0..7, and D[4*(l/16)+e][l%16] for e = 0..3.
//
// Features of a (cell, point) pair, all O(1): d/r (3, d = com - point),
// size/r (1), (com - cell centre)/size (3), Q/(M size^2) components xx, yy,
// xy, xz, yz (5).  Targets: R = (a_exact - a_multipole) r^2 / M / yscale_R
// and J = d R / d point r^3 / M / yscale_J.
//
// Parameters: FP32 master copy (W1[64][12] b1[64] W2[64][64] b2[64]
// W3[64][64] b3[64] W4[9][64] b4[9]) and BF16 copies padded for the MFMA
// operands (W1[64][32], W2[64][64], W3[64][64], W4[32][64]).
#include "bh_ctx.h"
#include "bh_mlp.h"

#include <algorithm>
#include <cmath>

namespace bh {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Activation: act(z) = z h(z) with the cubic sigmoid h = 1/2 + t (3/4 - t^2/4),
// t = clamp(z/3, -1, 1): h rises from 0 at z = -3 to 1 at z = 3 with SiLU's
// slope 1/4 at 0 and a continuous derivative (h' = (1 - t^2)/4).  Six
// full-rate vector instructions (multiply, med3, three FMA/multiplies, the
// product) where SiLU's exp and reciprocal run at a quarter of the rate.
// Training (forward and backward) and k_mlp_forward use this form.
__device__ __forceinline__ float act_h(float z, float &t)
{
    t = __builtin_amdgcn_fmed3f(z * (1.f / 3.f), -1.f, 1.f);
    return fmaf(t, fmaf(-0.25f, t * t, 0.75f), 0.5f);
}
__device__ __forceinline__ float act(float z)
{
    float t;
    return z * act_h(z, t);
}
__device__ __forceinline__ float act_grad(float z)
{
    float t;
    const float h = act_h(z, t);
    return fmaf(z * 0.25f, fmaf(-t, t, 1.f), h);
}
// The same polynomial as the smoothstep h = u^2 (3 - 2u) of u = (t + 1)/2 =
// clamp(z/6 + 1/2, 0, 1), whose clamp folds into the FMA's output clamp:
// five instructions.  The learned kernel's inference uses it; it rounds
// differently from act() at the last bit of h.
__device__ __forceinline__ float act_fast(float z)
{
    const float u = __builtin_amdgcn_fmed3f(fmaf(z, 1.f / 6.f, 0.5f), 0.f, 1.f);
    return z * ((u * u) * fmaf(-2.f, u, 3.f));
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// A operand of tile (ti, kk) from a row-major bf16 matrix (row stride lda),
// 16-byte aligned rows: rows ti*16.., columns kk..kk+31.
__device__ __forceinline__ bf16x8 load_rows(const bf16 *M, int lda, int r0, int k0)
{
    const int l = lane_id();
    return *reinterpret_cast<const bf16x8 *>(M + (r0 + (l & 15)) * lda + k0 + 8 * (l >> 4));
}

// A operand element-wise: A[i][k] = M[k*ld + i] (transposed access).
__device__ __forceinline__ bf16x8 load_cols(const bf16 *M, int ld, int r0, int k0)
{
    const int l = lane_id();
    bf16x8 v;
    const int i = r0 + (l & 15), kb = k0 + 8 * (l >> 4);
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] = M[(kb + e) * ld + i];
    return v;
}

// Weight offsets inside the BF16 parameter block.
// W4 is padded to 32 zero rows so that the transposed product in the backward
// pass (K = 32) stays inside the block.
constexpr int kW1b = 0, kW2b = kW1b + 64 * 32, kW3b = kW2b + 64 * 64, kW4b = kW3b + 64 * 64, kWbTotal = kW4b + 32 * 64;

struct Acts {
    bf16 *x;            // [64][32] inputs
    bf16 *z[3];         // pre-activations of the hidden layers [64][64] (training)
    bf16 *h[3];         // hidden activations [64][64]; may all alias (inference)
};

// One hidden layer: out[64][64] = act(in[64][K] W^T + b).  in/out may alias
// (all reads happen before the writes).  zsave != nullptr keeps Z (bf16).
template <int K>
__device__ void layer_hidden(const bf16 *in, const bf16 *W, const float *b, bf16 *out, bf16 *zsave)
{
    const int l = lane_id();
    f32x4 acc[4][4];
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int tj = 0; tj < 4; tj++) {
            acc[ti][tj] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < K; kk += 32)
                acc[ti][tj] = mfma(load_rows(in, K, ti * 16, kk), load_rows(W, K, tj * 16, kk), acc[ti][tj]);
        }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int tj = 0; tj < 4; tj++) {
            const int col = tj * 16 + (l & 15);
            const float bias = b[col];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int row = ti * 16 + 4 * (l >> 4) + e;
                const float z = acc[ti][tj][e] + bias;
                if (zsave) zsave[row * 64 + col] = (bf16)z;
                out[row * 64 + col] = (bf16)act(z);
            }
        }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Output layer: y[64][kMlpOut] (rows in lanes as D fragments) = h W4^T + b4.
// Returns for lane l the 4 rows 4*(l/16)+e of tile ti in column l%16.
__device__ void layer_out(const bf16 *h, const bf16 *W4, const float *b4, float y[4][4])
{
    const int l = lane_id();
#pragma unroll
    for (int ti = 0; ti < 4; ti++) {
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 64; kk += 32) acc = mfma(load_rows(h, 64, ti * 16, kk), load_rows(W4, 64, 0, kk), acc);
        const int col = l & 15;
        const float bias = col < kMlpOut ? b4[col] : 0.f;
#pragma unroll
        for (int e = 0; e < 4; e++) y[ti][e] = acc[e] + bias;
    }
}

// Forward pass of 64 rows already in a.x; y fragments as in layer_out.
__device__ void mlp_forward_wave(const Acts &a, const bf16 *Wb, const float *bias, float y[4][4])
{
    layer_hidden<32>(a.x, Wb + kW1b, bias + 0, a.h[0], a.z[0]);
    layer_hidden<64>(a.h[0], Wb + kW2b, bias + 64, a.h[1], a.z[1]);
    layer_hidden<64>(a.h[1], Wb + kW3b, bias + 128, a.h[2], a.z[2]);
    layer_out(a.h[2], Wb + kW4b, bias + 192, y);
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stage 64 rows of 12 FP32 features (rows >= nrow are zero) into a.x.
__device__ void stage_features(const Acts &a, const float *f, int nrow)
{
    const int l = lane_id();
    for (int c = 0; c < 32; c++) {
        float v = (l < nrow && c < kMlpIn) ? f[(int64_t)l * kMlpIn + c] : 0.f;
        a.x[l * 32 + c] = (bf16)v;
    }
    wave_sync_lds();
}

// ---- register-resident inference -------------------------------------------------
// The transposed product D' = W H^T (A = weight rows, B = activations) leaves
// in lane l the outputs j = 16t + 4(l/16) + e of sample l%16, which is the B
// operand layout of the next layer (k = 8(l/16) + e') once K is permuted
// inside each 32-wide step: slot e' of lane group q holds k = 32s + 4q + e'
// (e' < 4) or 32s + 16 + 4q + e' - 4 (e' >= 4).  The weight fragments are
// loaded once with that permutation and stay in VGPRs; activations never
// leave registers.  Rounding points are those of the LDS path: bf16 inputs,
// weights and hidden activations, FP32 accumulation and biases.
__device__ __forceinline__ int kperm(int s, int q, int e) { return e < 4 ? 32 * s + 4 * q + e : 32 * s + 12 + 4 * q + e; }

struct RegW {
    bf16x8 w1[4], w2[4][2], w3[4][2], w4[2];
};

__device__ __forceinline__ void load_regw(RegW &R, const bf16 *Wb)
{
    const int l = lane_id(), r = l & 15, q = l >> 4;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        R.w1[t] = *reinterpret_cast<const bf16x8 *>(Wb + kW1b + (16 * t + r) * 32 + 8 * q);
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int e = 0; e < 8; e++) {
                R.w2[t][s][e] = Wb[kW2b + (16 * t + r) * 64 + kperm(s, q, e)];
                R.w3[t][s][e] = Wb[kW3b + (16 * t + r) * 64 + kperm(s, q, e)];
            }
    }
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int e = 0; e < 8; e++) R.w4[s][e] = Wb[kW4b + r * 64 + kperm(s, q, e)];   // rows >= kMlpOut are zero
}


__device__ __forceinline__ void act_pack(const f32x4 h[4], bf16x8 hb[2])
{
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            hb[s][e] = (bf16)act_fast(h[2 * s][e]);
            hb[s][4 + e] = (bf16)act_fast(h[2 * s + 1][e]);
        }
}

__device__ __forceinline__ f32x4 bias4(const float *sb, int off)
{
    const float4 b = *reinterpret_cast<const float4 *>(sb + off);
    return f32x4{b.x, b.y, b.z, b.w};
}

// Forward of 16 samples (lane l: sample l%16) from the layer-1 B operand xb
// (inputs 8(l/16)..+7 of the sample); lane l returns outputs 4(l/16)..+3 of
// its sample (outputs >= kMlpOut are zero).
__device__ __forceinline__ f32x4 mlp_fwd_reg(const RegW &R, const float *sb, bf16x8 xb)
{
    const int q = lane_id() >> 4;
    f32x4 h[4];
    bf16x8 hb[2];
#pragma unroll
    for (int t = 0; t < 4; t++) h[t] = mfma(R.w1[t], xb, bias4(sb, 16 * t + 4 * q));
    act_pack(h, hb);
#pragma unroll
    for (int t = 0; t < 4; t++) h[t] = mfma(R.w2[t][1], hb[1], mfma(R.w2[t][0], hb[0], bias4(sb, 64 + 16 * t + 4 * q)));
    act_pack(h, hb);
#pragma unroll
    for (int t = 0; t < 4; t++) h[t] = mfma(R.w3[t][1], hb[1], mfma(R.w3[t][0], hb[0], bias4(sb, 128 + 16 * t + 4 * q)));
    act_pack(h, hb);
    return mfma(R.w4[1], hb[1], mfma(R.w4[0], hb[0], bias4(sb, 192 + 4 * q)));
}

// Layer-1 B operand of sub-tile sub from 64 staged rows of 16 bf16.
__device__ __forceinline__ bf16x8 xb_of(const bf16 *sx, int sub)
{
    const int l = lane_id(), r = 16 * sub + (l & 15), q = l >> 4;
    bf16x8 v;
    if (q < 2) {
        v = *reinterpret_cast<const bf16x8 *>(sx + r * 16 + 8 * q);
    } else {
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = (bf16)0.f;
    }
    return v;
}

__device__ __forceinline__ void stage_row(bf16 *sx, const float *f)
{
    const int l = lane_id();
    bf16x8 v0, v1;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        v0[k] = (bf16)f[k];
        v1[k] = (bf16)(k < kMlpIn - 8 ? f[8 + k] : 0.f);
    }
    reinterpret_cast<bf16x8 *>(sx + l * 16)[0] = v0;
    reinterpret_cast<bf16x8 *>(sx + l * 16)[1] = v1;
}

// ---- batch inference ---------------------------------------------------------


__global__ __launch_bounds__(64) void k_mlp_forward(const float *__restrict__ X, int64_t n, const bf16 *__restrict__ Wb,
                                                    const float *__restrict__ bias, float *__restrict__ Y)
{
    __shared__ __attribute__((aligned(16))) float sb[208];
    __shared__ __attribute__((aligned(16))) bf16 sx[64 * 16];
    const int l = lane_id();
    for (int k = l; k < 208; k += 64) sb[k] = bias[k];
    RegW R;
    load_regw(R, Wb);
    for (int64_t r0 = (int64_t)blockIdx.x * 64; r0 < n; r0 += (int64_t)gridDim.x * 64) {
        float f[kMlpIn];
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) f[k] = r0 + l < n ? X[(r0 + l) * kMlpIn + k] : 0.f;
        stage_row(sx, f);
        wave_sync_lds();
#pragma unroll
        for (int sub = 0; sub < 4; sub++) {
            const f32x4 y = mlp_fwd_reg(R, sb, xb_of(sx, sub));
            const int64_t row = r0 + 16 * sub + (l & 15);
            const int o = 4 * (l >> 4);
            if (row < n)
#pragma unroll
                for (int e = 0; e < 4; e++)
                    if (o + e < kMlpOut) Y[row * kMlpOut + o + e] = y[e];
        }
        wave_sync_lds();
    }
}

// ---- training: forward + backward, per-wave partial gradients ----------------

// dW[out][in] += sum_s dZ[s][out] * H[s][in]: A = dZ^T (transposed access of the
// [64][ldz] dZ block), B = H rows.  Tiles: out 16*ti.., in 16*tj.., K = 64 samples.
template <int OUT_T, int IN_T>
__device__ void grad_w(const bf16 *dZ, int ldz, const bf16 *H, int ldh, float *gW, int in_real, int out_real)
{
    const int l = lane_id();
#pragma unroll
    for (int ti = 0; ti < OUT_T; ti++)
#pragma unroll
        for (int tj = 0; tj < IN_T; tj++) {
            f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < 64; kk += 32) {
                bf16x8 av = load_cols(dZ, ldz, ti * 16, kk);
                bf16x8 bv;
                const int j = tj * 16 + (l & 15), kb = kk + 8 * (l >> 4);
#pragma unroll
                for (int e = 0; e < 8; e++) bv[e] = H[(kb + e) * ldh + j];
                acc = mfma(av, bv, acc);
            }
            const int in = tj * 16 + (l & 15);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int out = ti * 16 + 4 * (l >> 4) + e;
                if (out < out_real && in < in_real) gW[out * in_real + in] = acc[e];
            }
        }
}

// dH[s][in] = sum_out dZ[s][out] W[out][in]: A = dZ rows (K = out), B[k=out][j=in]
// = W[out][in] read element-wise from the BF16 weights (row stride ldw).
template <int KOUT>
__device__ void grad_h(const bf16 *dZ, int ldz, const bf16 *W, int ldw, float dH[4][4][4])
{
    const int l = lane_id();
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int tj = 0; tj < 4; tj++) {
            f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < KOUT; kk += 32) {
                bf16x8 av = load_rows(dZ, ldz, ti * 16, kk);
                bf16x8 bv;
                const int j = tj * 16 + (l & 15), kb = kk + 8 * (l >> 4);
#pragma unroll
                for (int e = 0; e < 8; e++) bv[e] = W[(kb + e) * ldw + j];
                acc = mfma(av, bv, acc);
            }
#pragma unroll
            for (int e = 0; e < 4; e++) dH[ti][tj][e] = acc[e];
        }
}

// Writes dZ = dH * act'(Z) (bf16) into dz[64][64] and db[64] partial sums.
__device__ void backprop_act(const float dH[4][4][4], const bf16 *Z, bf16 *dz, float *gb)
{
    const int l = lane_id();
    wave_sync_lds();
    float colsum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int tj = 0; tj < 4; tj++) {
            const int col = tj * 16 + (l & 15);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int row = ti * 16 + 4 * (l >> 4) + e;
                const float g = dH[ti][tj][e] * act_grad((float)Z[row * 64 + col]);
                const bf16 gb16 = (bf16)g;
                dz[row * 64 + col] = gb16;
                colsum[tj] += (float)gb16;
            }
        }
    // column sums over the 64 rows: lanes with the same l%16 hold the same columns
#pragma unroll
    for (int tj = 0; tj < 4; tj++) {
        float v = colsum[tj];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (l < 16) gb[tj * 16 + l] = v;
    }
    wave_sync_lds();
}

// One wave: 64 samples -> partial gradient (kMlpParams floats, FP32 master
// layout) and partial loss.  rows >= nrow contribute nothing.
__global__ __launch_bounds__(64) void k_mlp_grad(const float *__restrict__ X, const float *__restrict__ Yt, int64_t n,
                                                 float inv_batch, const bf16 *__restrict__ Wb,
                                                 const float *__restrict__ bias, float *__restrict__ gpart,
                                                 float *__restrict__ lpart)
{
    __shared__ __attribute__((aligned(16))) bf16 sx[64 * 32];
    __shared__ __attribute__((aligned(16))) bf16 sz[3][64 * 64];
    __shared__ __attribute__((aligned(16))) bf16 sh[3][64 * 64];
    __shared__ __attribute__((aligned(16))) bf16 sd[64 * 64];
    Acts a{sx, {sz[0], sz[1], sz[2]}, {sh[0], sh[1], sh[2]}};
    const int l = lane_id();
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const int nrow = (int)std::min<int64_t>(64, n - r0);
    stage_features(a, X + r0 * kMlpIn, nrow);
    float y[4][4];
    mlp_forward_wave(a, Wb, bias, y);
    float *g = gpart + (int64_t)blockIdx.x * kMlpParams;
    // dY = 2 (y - t) / batch on real rows/columns; kept as bf16 dZ4 [64][64] (cols < kMlpOut)
    const int col = l & 15;
    float lsum = 0.f, b4s = 0.f;
    for (int i = l; i < 64 * 64; i += 64) sd[i] = (bf16)0.f;
    wave_sync_lds();
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int row = ti * 16 + 4 * (l >> 4) + e;
            if (col < kMlpOut && row < nrow) {
                const float d = y[ti][e] - Yt[(r0 + row) * kMlpOut + col];
                lsum += d * d;
                const bf16 gd = (bf16)(2.f * d * inv_batch);
                sd[row * 64 + col] = gd;
                b4s += (float)gd;
            }
        }
    wave_sync_lds();
    // loss and db4
    for (int d = 32; d >= 1; d >>= 1) lsum += __shfl_xor(lsum, d, 64);
    if (l == 0) lpart[blockIdx.x] = lsum * inv_batch;
    b4s += __shfl_xor(b4s, 16, 64);
    b4s += __shfl_xor(b4s, 32, 64);
    if (l < kMlpOut) g[kOffB4 + l] = b4s;
    // layer 4 (W4 rows >= kMlpOut are zero, and so are dZ4's columns)
    grad_w<1, 4>(sd, 64, sh[2], 64, g + kOffW4, 64, kMlpOut);
    float dH[4][4][4];
    grad_h<32>(sd, 64, Wb + kW4b, 64, dH);
    // layer 3
    backprop_act(dH, sz[2], sd, g + kOffB3);
    grad_w<4, 4>(sd, 64, sh[1], 64, g + kOffW3, 64, 64);
    grad_h<64>(sd, 64, Wb + kW3b, 64, dH);
    // layer 2
    backprop_act(dH, sz[1], sd, g + kOffB2);
    grad_w<4, 4>(sd, 64, sh[0], 64, g + kOffW2, 64, 64);
    grad_h<64>(sd, 64, Wb + kW2b, 64, dH);
    // layer 1
    backprop_act(dH, sz[0], sd, g + kOffB1);
    grad_w<4, 1>(sd, 64, sx, 32, g + kOffW1, kMlpIn, 64);
}

// Sum of the per-wave partial gradients in wave order (deterministic).
__global__ void k_mlp_reduce(const float *__restrict__ gpart, int nw, float *__restrict__ grad, const float *lpart,
                             float *loss)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) {
        float s = 0.f;
        for (int w = 0; w < nw; w++) s += gpart[(int64_t)w * kMlpParams + p];
        grad[p] = s;
    }
    if (p == 0) {
        float s = 0.f;
        for (int w = 0; w < nw; w++) s += lpart[w];
        *loss = s;
    }
}

// Adam step on the FP32 master parameters, then refresh the BF16 copies.
__global__ void k_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ m, float *__restrict__ v,
                       float lr, float b1, float b2, float eps, float c1, float c2)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float g = G[p];
    float mm = b1 * m[p] + (1.f - b1) * g;
    float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}


// ---- features and the learned correction of the walk ---------------------------

// Features of cell c seen from point p (FP32, relative coordinates: rel = com -
// p).  Returns scale = M / r^2 (the unit of the residual; the Jacobian's is
// scale / r) or 0 if r == 0; *rinv = 1 / r.
__device__ __forceinline__ float cell_features(float rx, float ry, float rz, float4 A, float4 B, float4 F0, float4 F1,
                                               float *f, float *rinv)
{
    const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
    if (!(r2 > 0.f)) {
        for (int k = 0; k < kMlpIn; k++) f[k] = 0.f;
        *rinv = 0.f;
        return 0.f;
    }
    const float inv = rsqrtf(r2);
    *rinv = inv;
    f[0] = rx * inv;
    f[1] = ry * inv;
    f[2] = rz * inv;
    f[3] = sqrtf(B.w) * inv;   // size / r
    f[4] = F0.x;
    f[5] = F0.y;
    f[6] = F0.z;
    f[7] = F1.x;               // Qxx
    f[8] = F1.y;               // Qyy
    f[9] = F1.z;               // Qxy
    f[10] = F1.w;              // Qxz
    f[11] = F0.w;              // Qyz
    return A.w / r2;
}

// Flatten the recorded (group, cells) of every region into one entry list:
// entry e -> cell id and group; per group: first entry, count, centre.
__global__ void k_flatten(const int *__restrict__ lrec, int lcap, const uint32_t *__restrict__ off, int64_t ecap,
                          int *ecell, int *egrp, int *gstart, int *gcnt, float4 *gcentre)
{
    const int *reg = lrec + (int64_t)blockIdx.x * lcap;
    const int used = reg[0];
    int64_t e = off[blockIdx.x];
    for (int p = 1; p < used;) {
        const int g = reg[p];
        const int n = e + reg[p + 1] <= ecap ? reg[p + 1] : 0;
        if (threadIdx.x == 0) {
            gstart[g] = (int)e;
            gcnt[g] = n;
            gcentre[g] = make_float4(__int_as_float(reg[p + 2]), __int_as_float(reg[p + 3]),
                                     __int_as_float(reg[p + 4]), 0.f);
        }
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            ecell[e + k] = reg[p + 5 + k];
            egrp[e + k] = g;
        }
        e += n;
        p += 5 + reg[p + 1];
    }
}

__global__ void k_zero_i32n(int *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_learned_count(const uint32_t *total, int64_t ecap, unsigned long long *stat)
{
    stat[S_NMLP] += *total;
    if ((int64_t)*total > ecap) stat[S_FLAGS] |= (unsigned long long)F_ECAP_OVF;
}

// Adds the correction c (uniform per group) to the accelerations of group
// g's members.  The recording walk left each member's (a, phi) in sorted
// order in sacc; the member's final values (plus the correction) go to the
// input-order arrays in one write.
__device__ __forceinline__ void apply_group(int g, float4 c, const int *__restrict__ gfirst,
                                            const int *__restrict__ gcount, const uint32_t *__restrict__ tperm,
                                            const double4 *__restrict__ sacc, double *phi, double *acc)
{
    const int f = gfirst[g], n = gcount[g];
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int64_t o = tperm[f + k];
        const double4 r = sacc[f + k];
        acc[3 * o] = r.x + (double)c.x;
        acc[3 * o + 1] = r.y + (double)c.y;
        acc[3 * o + 2] = r.z + (double)c.z;
        phi[o] = r.w;
    }
}

// First-order correction of the members [f, f + n) of a group: R + J (x -
// centre), with the member's position x from its FP32 hi/lo parts and J
// symmetric (j = xx xy xz yy yz zz); the recording walk left each member's
// (a, phi) in sorted order in sacc, and the final values go to the
// input-order arrays in one write.  Two rows of 64 per round, both rows'
// loads issued before either row's stores.
__device__ __forceinline__ void apply_range1(int f, int n, const float *R, const float *j, float4 ctr,
                                             const uint32_t *__restrict__ tperm, const float4 *__restrict__ thi,
                                             const float4 *__restrict__ tlo, const double4 *__restrict__ sacc,
                                             double *phi, double *acc)
{
    const int l = threadIdx.x & 63;
    for (int k0 = 0; k0 < n; k0 += 128) {
        float4 h[2], lo[2];
        double4 r[2];
        int64_t o[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int k = k0 + 64 * u + l;
            if (k < n) {
                h[u] = thi[f + k];
                lo[u] = tlo[f + k];
                o[u] = tperm[f + k];
                r[u] = sacc[f + k];
            }
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int k = k0 + 64 * u + l;
            if (k < n) {
                const float dx = (h[u].x - ctr.x) + lo[u].x, dy = (h[u].y - ctr.y) + lo[u].y,
                            dz = (h[u].z - ctr.z) + lo[u].z;
                acc[3 * o[u]] = r[u].x + (double)fmaf(j[0], dx, fmaf(j[1], dy, fmaf(j[2], dz, R[0])));
                acc[3 * o[u] + 1] = r[u].y + (double)fmaf(j[1], dx, fmaf(j[3], dy, fmaf(j[4], dz, R[1])));
                acc[3 * o[u] + 2] = r[u].z + (double)fmaf(j[2], dx, fmaf(j[4], dy, fmaf(j[5], dz, R[2])));
                phi[o[u]] = r[u].w;
            }
        }
    }
}

constexpr int kLearnedWavesPerCU = 8;    // k_learned_group's grid (4 to 16 measured alike)

// The learned correction of whole groups, one wave per group: the group's
// entries (contiguous, flattened) through the MLP 64 at a time, each entry's
// prediction written to res when it is kept (per entry: R xyz, J xx | J xy xz
// yy yz | J zz),
// summed per output by the lanes that hold it, then R + J (x - centre) added
// to every member (apply_range1).  The members' scattered stores overlap the
// other waves' MFMA work; the sums are in a fixed order (deterministic).
__global__ __launch_bounds__(64) void k_learned_group(const int *__restrict__ ecell, const int *__restrict__ gstart,
                                                      const int *__restrict__ gcnt,
                                                      const unsigned long long *__restrict__ stat,
                                                      const float4 *__restrict__ gcentre, const float4 *__restrict__ cA,
                                                      const float4 *__restrict__ cB, const float4 *__restrict__ cF0,
                                                      const float4 *__restrict__ cF1, const bf16 *__restrict__ Wb,
                                                      const float *__restrict__ bias, float yscale, float jscale,
                                                      float4 *__restrict__ res, const int *__restrict__ gfirst,
                                                      const int *__restrict__ gcount, const uint32_t *__restrict__ tperm,
                                                      const float4 *__restrict__ thi, const float4 *__restrict__ tlo,
                                                      const double4 *__restrict__ sacc, double *phi, double *acc)
{
    __shared__ __attribute__((aligned(16))) float sb[208];
    __shared__ __attribute__((aligned(16))) bf16 sx[64 * 16];
    const int l = lane_id(), q = l >> 4;
    for (int k = l; k < 208; k += 64) sb[k] = bias[k];
    RegW R;
    load_regw(R, Wb);
    // the weights in place before the loop, so the loop's own waits (the
    // weights' first uses inside it) do not also wait for its prefetches
    __builtin_amdgcn_s_waitcnt(0);
    const int ng = (int)stat[S_NGROUPS];
    // Loads run one step ahead of their use: the next group's entry range
    // while this group runs, the next chunk's cell ids before this chunk's
    // MLP and their cell records after its first sub-tile (a group's last
    // chunk fetches the next group's first one), so the dependent loads
    // overlap the MFMA work instead of stalling the wave.  The records of the
    // chunk being processed are only read for its features, before the
    // sub-tiles.
    int g = blockIdx.x;
    int s = 0, n = 0;
    if (g < ng) {
        s = gstart[g];
        n = gcnt[g];
    }
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 A = zero4, B = A, F0 = A, F1 = A;
    bool pre = false;   // A..F1 hold this group's first chunk
    float4 ctrL = g < ng ? gcentre[g] : zero4;
    for (; g < ng; g += gridDim.x) {
        const int gn = g + gridDim.x;
        int sN = 0, nN = 0;
        float4 ctrN = zero4;
        if (gn < ng) {
            sN = gstart[gn];
            nN = gcnt[gn];
            ctrN = gcentre[gn];
        }
        // (a group without entries adds zero: R = J = 0, and its centre,
        // which it may not have, is replaced by 0 so that J dx stays 0)
        const float4 ctr = n > 0 ? ctrL : zero4;
        const int mf = gfirst[g], mn = gcount[g];
        if (!pre && n > 0) {
            const unsigned c = (unsigned)ecell[s + min(l, n - 1)];
            A = cA[c];
            B = cB[c];
            F0 = cF0[c];
            F1 = cF1[c];
        }
        float a4[4] = {0.f, 0.f, 0.f, 0.f};
        for (int t0 = 0; t0 < n; t0 += 64) {
            float f[kMlpIn];
            float scale = 0.f, rinv = 0.f;
            if (t0 + l < n) {
                scale = cell_features((A.x - ctr.x) + B.x, (A.y - ctr.y) + B.y, (A.z - ctr.z) + B.z, A, B, F0, F1, f,
                                      &rinv);
            } else {
#pragma unroll
                for (int k = 0; k < kMlpIn; k++) f[k] = 0.f;
            }
            stage_row(sx, f);
            wave_sync_lds();
            // (unconditional, with wave-uniform choices: a lane past the end
            // reads the last entry's cell, and without a next chunk or group
            // every lane does, so no per-lane test and no merge of registers
            // waits on these loads; unsigned, so forming the address does not
            // wait on the id either)
            const int ie = t0 + 64 < n ? s + min(t0 + 64 + l, n - 1) : (nN > 0 ? sN + min(l, nN - 1) : s + n - 1);
            const unsigned cN = (unsigned)ecell[ie];
#pragma unroll
            for (int sub = 0; sub < 4; sub++) {
                if (t0 + 16 * sub >= n) break;
                // lane l: outputs 4q..4q+3 of entry 16 sub + l % 16
                const f32x4 y = mlp_fwd_reg(R, sb, xb_of(sx, sub));
                if (sub == 0) {
                    // (kept after this sub-tile's MFMAs: the compiler would
                    // otherwise hoist the address arithmetic, and its wait for
                    // the cell id, above them; the empty asm pins the id here)
                    __builtin_amdgcn_sched_barrier(0);
                    unsigned cq = cN;
                    asm volatile("" : "+v"(cq));
                    A = cA[cq];
                    B = cB[cq];
                    F0 = cF0[cq];
                    F1 = cF1[cq];
                }
                const float sc = __shfl(scale, 16 * sub + (l & 15), 64);
                const float ri = __shfl(rinv, 16 * sub + (l & 15), 64);
                const int er = t0 + 16 * sub + (l & 15);
                if (q < 3 && er < n) {
                    const float wr = yscale * sc, wj = jscale * sc * ri;
                    float4 o;
                    if (q == 0) o = make_float4(y[0] * wr, y[1] * wr, y[2] * wr, y[3] * wj);
                    else if (q == 1) o = make_float4(y[0] * wj, y[1] * wj, y[2] * wj, y[3] * wj);
                    else o = make_float4(y[0] * wj, 0.f, 0.f, 0.f);
                    if (res) res[3 * (int64_t)(s + er) + q] = o;
                    a4[0] += o.x;
                    a4[1] += o.y;
                    a4[2] += o.z;
                    a4[3] += o.w;
                }
            }
            wave_sync_lds();
        }
        // sums over the 16 lanes of each quarter, then the nine outputs
#pragma unroll
        for (int m = 1; m < 16; m <<= 1)
#pragma unroll
            for (int k = 0; k < 4; k++) a4[k] += __shfl_xor(a4[k], m, 64);
        float v[9];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = __shfl(a4[k], 0, 64);
            v[4 + k] = __shfl(a4[k], 16, 64);
        }
        v[8] = __shfl(a4[0], 32, 64);
        apply_range1(mf, mn, v, v + 3, ctr, tperm, thi, tlo, sacc, phi, acc);
        pre = n > 0 && nN > 0;
        s = sN;
        n = nN;
        ctrL = ctrN;
    }
}

// Adds the per-group corrections corr[] to the group members' accelerations.
__global__ void k_apply_corr(const int *__restrict__ gfirst, const int *__restrict__ gcount,
                             const unsigned long long *__restrict__ stat, const uint32_t *__restrict__ tperm,
                             const float4 *__restrict__ corr, const double4 *__restrict__ sacc, double *phi,
                             double *acc)
{
    const int ng = (int)stat[S_NGROUPS];
    for (int g = blockIdx.x; g < ng; g += gridDim.x) apply_group(g, corr[g], gfirst, gcount, tperm, sacc, phi, acc);
}

__global__ void k_zero_f4(float4 *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Cells recorded per group (sample mode): every walked group has one record
// in the region of the wave that walked it; groups without one keep 0.
__global__ void k_group_cells(const int *__restrict__ lrec, int lcap, int nreg, int ng, uint32_t *cnt)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nreg) return;
    const int *reg = lrec + (int64_t)r * lcap;
    for (int p = 1; p < reg[0]; p += 5 + reg[p + 1])
        if ((unsigned)reg[p] < (unsigned)ng) cnt[reg[p]] = (uint32_t)reg[p + 1];
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Cells recorded per region (sample mode), for the sample offsets.
__global__ void k_region_cells(const int *__restrict__ lrec, int lcap, int nreg, uint32_t *cnt)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nreg) return;
    const int *reg = lrec + (int64_t)r * lcap;
    uint32_t c = 0;
    for (int p = 1; p < reg[0]; p += 5 + reg[p + 1]) c += reg[p + 1];
    cnt[r] = c;
}

// Exact residual (a_exact - a_multipole) of cell c at point t, FP64; the wave
// shares the direct sum over the cell's bodies; valid in lane 0.
__device__ void exact_residual(int c, double tx, double ty, double tz, const int2 *__restrict__ fnb,
                               const double *__restrict__ mo, const double4 *__restrict__ bd, double e2, int usequad,
                               double *res)
{
    const int l = lane_id();
    const int2 fn = fnb[c];
    double ex = 0.0, ey = 0.0, ez = 0.0;
    for (int b = l; b < fn.y; b += 64) {
        const double4 q = bd[fn.x + b];
        const double dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
        const double r2 = dx * dx + dy * dy + dz * dz + e2;
        const double w = q.w / (r2 * sqrt(r2));
        ex += dx * w;
        ey += dy * w;
        ez += dz * w;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        ex += __shfl_xor(ex, d, 64);
        ey += __shfl_xor(ey, d, 64);
        ez += __shfl_xor(ez, d, 64);
    }
    const double *m = mo + 10 * (int64_t)c;
    const double dx = m[1] - tx, dy = m[2] - ty, dz = m[3] - tz;
    const double r2 = dx * dx + dy * dy + dz * dz + e2;
    const double ri = 1.0 / sqrt(r2), ri2 = ri * ri;
    const double mr3 = m[0] * ri * ri2;
    double mx = dx * mr3, my = dy * mr3, mz = dz * mr3;
    if (usequad) {
        const double *Q = m + 4;   // xx xy xz yy yz zz
        const double qx = Q[0] * dx + Q[1] * dy + Q[2] * dz;
        const double qy = Q[1] * dx + Q[3] * dy + Q[4] * dz;
        const double qz = Q[2] * dx + Q[4] * dy + Q[5] * dz;
        const double r5 = ri2 * ri2 * ri;
        const double phq = -0.5 * r5 * (dx * qx + dy * qy + dz * qz);
        const double cf = 5.0 * phq * ri2;
        mx -= dx * cf + qx * r5;
        my -= dy * cf + qy * r5;
        mz -= dz * cf + qz * r5;
    }
    res[0] = ex - mx;
    res[1] = ey - my;
    res[2] = ez - mz;
}

// Diagnostic: the correction a perfect model would give (exact residual at the
// group centre for every recorded cell).
__global__ __launch_bounds__(64) void k_exact_corr(const int *__restrict__ lrec, int lcap, const int2 *__restrict__ fnb,
                                                   const double *__restrict__ mo, const double4 *__restrict__ bd,
                                                   double eps, int usequad, float4 *corr)
{
    const int *reg = lrec + (int64_t)blockIdx.x * lcap;
    const int used = reg[0];
    for (int p = 1; p < used;) {
        const int g = reg[p], n = reg[p + 1];
        const double tx = __int_as_float(reg[p + 2]), ty = __int_as_float(reg[p + 3]), tz = __int_as_float(reg[p + 4]);
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int k = 0; k < n; k++) {
            double r[3];
            exact_residual(reg[p + 5 + k], tx, ty, tz, fnb, mo, bd, eps * eps, usequad, r);
            ax += r[0];
            ay += r[1];
            az += r[2];
        }
        if (lane_id() == 0) corr[g] = make_float4((float)ax, (float)ay, (float)az, 0.f);
        p += 5 + n;
    }
}

// Residual field of cell c near point t, FP64: the residual (a_exact -
// a_multipole) at t and its Jacobian J[i][k] = d res_i / d t_k (row major, 9).
// The exact part is differentiated analytically (m (3 d d^T / r^5 - I / r^3)
// per body, d = body - t, softened r); the multipole part by central
// differences of the multipole formula with step 1e-4 r (FP64: the
// truncation error is ~1e-8 relative).  Valid in lane 0.
__device__ void residual_field(int c, double tx, double ty, double tz, const int2 *__restrict__ fnb,
                               const double *__restrict__ mo, const double4 *__restrict__ bd, double e2, int usequad,
                               double *res, double *J)
{
    const int l = lane_id();
    const int2 fn = fnb[c];
    double ex = 0.0, ey = 0.0, ez = 0.0;
    double j[6] = {0, 0, 0, 0, 0, 0};   // xx xy xz yy yz zz
    for (int b = l; b < fn.y; b += 64) {
        const double4 q = bd[fn.x + b];
        const double dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
        const double r2 = dx * dx + dy * dy + dz * dz + e2;
        const double ri = 1.0 / sqrt(r2);
        const double w = q.w * ri * ri * ri;
        const double w5 = 3.0 * w * ri * ri;
        ex += dx * w;
        ey += dy * w;
        ez += dz * w;
        j[0] += w5 * dx * dx - w;
        j[1] += w5 * dx * dy;
        j[2] += w5 * dx * dz;
        j[3] += w5 * dy * dy - w;
        j[4] += w5 * dy * dz;
        j[5] += w5 * dz * dz - w;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        ex += __shfl_xor(ex, d, 64);
        ey += __shfl_xor(ey, d, 64);
        ez += __shfl_xor(ez, d, 64);
#pragma unroll
        for (int k = 0; k < 6; k++) j[k] += __shfl_xor(j[k], d, 64);
    }
    const double *m = mo + 10 * (int64_t)c;
    auto multipole = [&](double px, double py, double pz, double *a) {
        const double dx = m[1] - px, dy = m[2] - py, dz = m[3] - pz;
        const double r2 = dx * dx + dy * dy + dz * dz + e2;
        const double ri = 1.0 / sqrt(r2), ri2 = ri * ri;
        const double mr3 = m[0] * ri * ri2;
        a[0] = dx * mr3;
        a[1] = dy * mr3;
        a[2] = dz * mr3;
        if (usequad) {
            const double *Q = m + 4;
            const double qx = Q[0] * dx + Q[1] * dy + Q[2] * dz;
            const double qy = Q[1] * dx + Q[3] * dy + Q[4] * dz;
            const double qz = Q[2] * dx + Q[4] * dy + Q[5] * dz;
            const double r5 = ri2 * ri2 * ri;
            const double phq = -0.5 * r5 * (dx * qx + dy * qy + dz * qz);
            const double cf = 5.0 * phq * ri2;
            a[0] -= dx * cf + qx * r5;
            a[1] -= dy * cf + qy * r5;
            a[2] -= dz * cf + qz * r5;
        }
    };
    double a0[3];
    multipole(tx, ty, tz, a0);
    res[0] = ex - a0[0];
    res[1] = ey - a0[1];
    res[2] = ez - a0[2];
    const double r = sqrt((m[1] - tx) * (m[1] - tx) + (m[2] - ty) * (m[2] - ty) + (m[3] - tz) * (m[3] - tz) + e2);
    const double h = 1e-4 * r;
    const double je[9] = {j[0], j[1], j[2], j[1], j[3], j[4], j[2], j[4], j[5]};
    for (int k = 0; k < 3; k++) {
        double ap[3], am[3];
        multipole(tx + (k == 0 ? h : 0.0), ty + (k == 1 ? h : 0.0), tz + (k == 2 ? h : 0.0), ap);
        multipole(tx - (k == 0 ? h : 0.0), ty - (k == 1 ? h : 0.0), tz - (k == 2 ? h : 0.0), am);
        for (int i = 0; i < 3; i++) J[3 * i + k] = je[3 * i + k] - (ap[i] - am[i]) / (2.0 * h);
    }
}

// Diagnostic: the first-order correction a perfect model of the residual
// field would give: per group the sums of the exact residuals and of their
// Jacobians at the group centre over the recorded cells (12 floats: R, J row
// major), and the centre.
__global__ __launch_bounds__(64) void k_exact_corr1(const int *__restrict__ lrec, int lcap,
                                                    const int2 *__restrict__ fnb, const double *__restrict__ mo,
                                                    const double4 *__restrict__ bd, double eps, int usequad,
                                                    float4 *corr)
{
    const int *reg = lrec + (int64_t)blockIdx.x * lcap;
    const int used = reg[0];
    for (int p = 1; p < used;) {
        const int g = reg[p], n = reg[p + 1];
        const float cx = __int_as_float(reg[p + 2]), cy = __int_as_float(reg[p + 3]), cz = __int_as_float(reg[p + 4]);
        double s[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < n; k++) {
            double r[3], J[9];
            residual_field(reg[p + 5 + k], cx, cy, cz, fnb, mo, bd, eps * eps, usequad, r, J);
            for (int t = 0; t < 3; t++) s[t] += r[t];
            for (int t = 0; t < 9; t++) s[3 + t] += J[t];
        }
        if (lane_id() == 0) {
            float4 *o = corr + 4 * (int64_t)g;
            o[0] = make_float4((float)s[0], (float)s[1], (float)s[2], cx);
            o[1] = make_float4((float)s[3], (float)s[4], (float)s[5], cy);
            o[2] = make_float4((float)s[6], (float)s[7], (float)s[8], cz);
            o[3] = make_float4((float)s[9], (float)s[10], (float)s[11], 0.f);
        }
        p += 5 + n;
    }
}

// Adds R + J (x - centre) per member (first-order correction).
__global__ void k_apply_corr1(const int *__restrict__ gfirst, const int *__restrict__ gcount,
                              const unsigned long long *__restrict__ stat, const uint32_t *__restrict__ tperm,
                              const float4 *__restrict__ thi, const float4 *__restrict__ tlo,
                              const float4 *__restrict__ corr, const double4 *__restrict__ sacc, double *phi,
                              double *acc)
{
    const int ng = (int)stat[S_NGROUPS];
    for (int g = blockIdx.x; g < ng; g += gridDim.x) {
        const float4 c0 = corr[4 * (int64_t)g], c1 = corr[4 * (int64_t)g + 1], c2 = corr[4 * (int64_t)g + 2],
                     c3 = corr[4 * (int64_t)g + 3];
        const float J[9] = {c1.x, c1.y, c1.z, c2.x, c2.y, c2.z, c3.x, c3.y, c3.z};
        const int f = gfirst[g], n = gcount[g];
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const float4 h = thi[f + k], lo = tlo[f + k];
            const float dx = (h.x - c0.w) + lo.x, dy = (h.y - c1.w) + lo.y, dz = (h.z - c2.w) + lo.z;
            const int64_t o = tperm[f + k];
            const double4 r = sacc[f + k];
            acc[3 * o] = r.x + (double)(c0.x + J[0] * dx + J[1] * dy + J[2] * dz);
            acc[3 * o + 1] = r.y + (double)(c0.y + J[3] * dx + J[4] * dy + J[5] * dz);
            acc[3 * o + 2] = r.z + (double)(c0.z + J[6] * dx + J[7] * dy + J[8] * dz);
            phi[o] = r.w;
        }
    }
}

// Training samples: for every recorded (group centre, cell), the features and
// the exact residual (FP64 softened direct sum over the cell's bodies minus
// the FP64 multipole force of the walk's formula) in units of M / r^2, and
// its Jacobian in units of M / r^3.
// Group g's samples go to goff[g] onwards, so the set is in group order
// whichever wave walked which group (training is then reproducible).
__global__ __launch_bounds__(64) void k_samples(const int *__restrict__ lrec, int lcap, const uint32_t *__restrict__ goff,
                                                int ng,
                                                const float4 *__restrict__ cA, const float4 *__restrict__ cB,
                                                const float4 *__restrict__ cF0, const float4 *__restrict__ cF1,
                                                const int2 *__restrict__ fnb, const double *__restrict__ mo,
                                                const double4 *__restrict__ bd, double eps, int usequad,
                                                float *__restrict__ X, float *__restrict__ Y,
                                                float4 *__restrict__ meta)
{
    const int l = lane_id();
    const int *reg = lrec + (int64_t)blockIdx.x * lcap;
    const int used = reg[0];
    const double e2 = eps * eps;
    for (int p = 1; p < used;) {
        const int n = reg[p + 1];
        if ((unsigned)reg[p] >= (unsigned)ng) {   // not counted by k_group_cells
            p += 5 + n;
            continue;
        }
        int64_t j = goff[reg[p]];
        const float cx = __int_as_float(reg[p + 2]), cy = __int_as_float(reg[p + 3]), cz = __int_as_float(reg[p + 4]);
        for (int k = 0; k < n; k++, j++) {
            const int c = reg[p + 5 + k];
            double r[3], J[9];
            residual_field(c, cx, cy, cz, fnb, mo, bd, e2, usequad, r, J);
            if (l == 0) {
                float f[kMlpIn], rinv;
                const float4 A = cA[c], B = cB[c];
                const float scale = cell_features((A.x - cx) + B.x, (A.y - cy) + B.y, (A.z - cz) + B.z, A, B, cF0[c],
                                                  cF1[c], f, &rinv);
                for (int t = 0; t < kMlpIn; t++) X[j * kMlpIn + t] = f[t];
                meta[j] = make_float4(cx, cy, cz, __int_as_float(c));
                const double inv = scale > 0.f ? 1.0 / (double)scale : 0.0;
                const double invj = scale > 0.f ? inv / (double)rinv : 0.0;
                float *y = Y + j * kMlpOut;
                y[0] = (float)(r[0] * inv);
                y[1] = (float)(r[1] * inv);
                y[2] = (float)(r[2] * inv);
                // the symmetric part of the Jacobian (the exact one is symmetric;
                // the average removes the difference-quotient noise)
                y[3] = (float)(J[0] * invj);
                y[4] = (float)(0.5 * (J[1] + J[3]) * invj);
                y[5] = (float)(0.5 * (J[2] + J[6]) * invj);
                y[6] = (float)(J[4] * invj);
                y[7] = (float)(0.5 * (J[5] + J[7]) * invj);
                y[8] = (float)(J[8] * invj);
            }
        }
        p += 5 + n;
    }
}

}  // namespace

// Training batch b of an epoch: rows (a*i + c) mod n (a coprime with n), targets
// divided by their block's scale and clipped to +-kClip.
constexpr float kClip = 10.f;
__global__ void k_gather_batch(const float *__restrict__ X, const float *__restrict__ Y, int64_t n, uint64_t a,
                               uint64_t c, int64_t first, int64_t cnt, float inv_yscale, float inv_jscale, float *bx,
                               float *by)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const uint64_t r = (uint64_t)((a * (uint64_t)(first + i) + c) % (uint64_t)n);
    for (int k = 0; k < kMlpIn; k++) bx[i * kMlpIn + k] = X[r * kMlpIn + k];
    for (int k = 0; k < kMlpOut; k++)
        by[i * kMlpOut + k] = fminf(fmaxf(Y[r * kMlpOut + k] * (k < kMlpOutR ? inv_yscale : inv_jscale), -kClip), kClip);
}

// Converts the FP32 master parameters to the padded BF16 MFMA layout and the
// FP32 bias block (b1 | b2 | b3 | b4 padded to 16).
__global__ void k_mlp_pack(const float *__restrict__ P, bf16 *__restrict__ Wb, float *__restrict__ bias)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < kWbTotal) {
        float v = 0.f;
        if (t < kW2b) {
            int o = t / 32, i = t % 32;
            v = i < kMlpIn ? P[kOffW1 + o * kMlpIn + i] : 0.f;
        } else if (t < kW3b) {
            v = P[kOffW2 + (t - kW2b)];
        } else if (t < kW4b) {
            v = P[kOffW3 + (t - kW3b)];
        } else {
            int o = (t - kW4b) / 64, i = (t - kW4b) % 64;
            v = o < kMlpOut ? P[kOffW4 + o * 64 + i] : 0.f;
        }
        Wb[t] = (bf16)v;
    }
    if (t < 64) {
        bias[t] = P[kOffB1 + t];
        bias[64 + t] = P[kOffB2 + t];
        bias[128 + t] = P[kOffB3 + t];
    }
    if (t < 16) bias[192 + t] = t < kMlpOut ? P[kOffB4 + t] : 0.f;
}

// ---- host side ---------------------------------------------------------------

void mlp_pack(bh_ctx *ctx)
{
    Mlp &M = ctx->mlp;
    k_mlp_pack<<<(kWbTotal + 255) / 256, 256, 0, ctx->stream>>>(ptr<float>(M.params), ptr<bf16>(M.wb), ptr<float>(M.bias));
    BH_HIP(hipGetLastError());
}

void mlp_set(bh_ctx *ctx, const float *host_params, const float *norm)
{
    Mlp &M = ctx->mlp;
    float *P = ensure<float>(M.params, kMlpParams);
    ensure<bf16>(M.wb, kWbTotal);
    ensure<float>(M.bias, 208);
    BH_HIP(hipMemcpyAsync(P, host_params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    if (norm) {
        for (int k = 0; k < kMlpNorm; k++) M.norm[k] = norm[k];
    }
    mlp_pack(ctx);
    BH_HIP(hipStreamSynchronize(ctx->stream));
    M.ready = true;
}


// Learned correction after a walk with lmode = 1: per-group corrections from
// the recorded cells, added to the members' accelerations.
void mlp_correct(bh_ctx *ctx, const bh_params &p)
{
    const bool exact = p.learned == 2;   // diagnostic: exact residual at the group centre
    const bool exact1 = p.learned == 3;  // diagnostic: exact residual and Jacobian, first order per member
    Mlp &M = ctx->mlp;
    const Tree &T = ctx->tree;
    if (T.nb == 0) return;   // the walk wrote zero forces directly
    hipStream_t st = ctx->stream;
    const int64_t ngcap = ctx->ntest;
    unsigned long long *stat = ptr<unsigned long long>(ctx->counters);
    const int gridg = (int)std::min<int64_t>(ngcap, (int64_t)ctx->ncu * 32);
    if (exact1) {
        float4 *corr = ensure<float4>(M.corr, 4 * ngcap);
        k_zero_f4<<<(int)((4 * ngcap + 255) / 256), 256, 0, st>>>(corr, (int)(4 * ngcap));
        k_exact_corr1<<<ctx->lnwave, 64, 0, st>>>(ptr<int>(ctx->lrec), (int)ctx->lcap, T.fnb, T.mo, T.bd, p.eps,
                                                  p.usequad, corr);
        k_apply_corr1<<<gridg, 64, 0, st>>>(ptr<int>(ctx->gfirst), ptr<int>(ctx->gcount), stat, T.tperm, T.thi, T.tlo,
                                            corr, ptr<double4>(ctx->sacc), ptr<double>(ctx->phi),
                                            ptr<double>(ctx->acc));
    } else if (exact) {
        float4 *corr = ensure<float4>(M.corr, ngcap);
        k_zero_f4<<<(int)((ngcap + 255) / 256), 256, 0, st>>>(corr, (int)ngcap);
        k_exact_corr<<<ctx->lnwave, 64, 0, st>>>(ptr<int>(ctx->lrec), (int)ctx->lcap, T.fnb, T.mo, T.bd, p.eps,
                                                 p.usequad, corr);
        k_apply_corr<<<gridg, 64, 0, st>>>(ptr<int>(ctx->gfirst), ptr<int>(ctx->gcount), stat, T.tperm, corr,
                                           ptr<double4>(ctx->sacc), ptr<double>(ctx->phi), ptr<double>(ctx->acc));
    } else {
        // flatten the records, one MLP tile per wave, per-group sums
        const int nreg = ctx->lnwave;
        uint32_t *cnt = ensure<uint32_t>(M.soff, 2 * (size_t)nreg + 2);
        uint32_t *off = cnt + nreg + 1;
        k_region_cells<<<(nreg + 255) / 256, 256, 0, st>>>(ptr<int>(ctx->lrec), (int)ctx->lcap, nreg, cnt);
        exclusive_scan_u32(ctx, cnt, off, nreg);
        if (M.ecap == 0) M.ecap = std::max<int64_t>(1 << 20, 16 * ctx->ntest);
        const int64_t ecap = M.ecap;
        int *ecell = ensure<int>(M.ecell, ecap), *egrp = ensure<int>(M.egrp, ecap);
        int *gstart = ensure<int>(M.gstart, ngcap), *gcnt = ensure<int>(M.gcnt, ngcap);
        float4 *gctr = ensure<float4>(M.gcentre, ngcap);
        float4 *res = ensure<float4>(M.res, 3 * ecap);
        k_zero_i32n<<<(int)((ngcap + 255) / 256), 256, 0, st>>>(gcnt, (int)ngcap);
        k_flatten<<<nreg, 64, 0, st>>>(ptr<int>(ctx->lrec), (int)ctx->lcap, off, ecap, ecell, egrp, gstart, gcnt, gctr);
        // the number of entries is only known on the device: launch over the
        // bound and let waves past the end exit
        k_learned_count<<<1, 1, 0, st>>>(off + nreg, ecap, stat);
        if (!ctx->capturing) BH_HIP(hipEventRecord(ctx->ev[8], st));
        k_learned_group<<<(int)std::min<int64_t>(ngcap, (int64_t)ctx->ncu * kLearnedWavesPerCU), 64, 0, st>>>(
            ecell, gstart, gcnt, stat, gctr, T.cA, T.cB, T.cF0, T.cF1, ptr<bf16>(M.wb), ptr<float>(M.bias),
            M.norm[0], M.norm[3], M.keep_res ? res : nullptr, ptr<int>(ctx->gfirst), ptr<int>(ctx->gcount), T.tperm, T.thi, T.tlo,
            ptr<double4>(ctx->sacc), ptr<double>(ctx->phi), ptr<double>(ctx->acc));
        if (!ctx->capturing) BH_HIP(hipEventRecord(ctx->ev[9], st));
    }
    BH_HIP(hipGetLastError());
}

// Training samples from the last sample-mode walk, appended to the set.
int64_t mlp_collect(bh_ctx *ctx, const bh_params &p)
{
    Mlp &M = ctx->mlp;
    const Tree &T = ctx->tree;
    hipStream_t st = ctx->stream;
    const int nreg = ctx->lnwave;
    const int ng = ctx->ngroups;
    if (ng <= 0) return 0;
    // samples in group order: per-group counts, then their exclusive scan
    uint32_t *cnt = ensure<uint32_t>(M.soff, 2 * (size_t)ng + 2);
    uint32_t *goff = cnt + ng + 1;
    k_fill_u32<<<(ng + 255) / 256, 256, 0, st>>>(cnt, ng, 0u);
    k_group_cells<<<(nreg + 255) / 256, 256, 0, st>>>(ptr<int>(ctx->lrec), (int)ctx->lcap, nreg, ng, cnt);
    exclusive_scan_u32(ctx, cnt, goff, ng);
    uint32_t total = 0;
    BH_HIP(hipMemcpyAsync(&total, goff + ng, sizeof total, hipMemcpyDeviceToHost, st));
    BH_HIP(hipStreamSynchronize(st));
    if (total == 0) return 0;
    float *nx = ensure<float>(M.snew_x, (size_t)total * kMlpIn), *ny = ensure<float>(M.snew_y, (size_t)total * kMlpOut);
    float4 *meta = ensure<float4>(M.smeta, total);
    k_samples<<<nreg, 64, 0, st>>>(ptr<int>(ctx->lrec), (int)ctx->lcap, goff, ng, T.cA, T.cB, T.cF0, T.cF1, T.fnb, T.mo,
                                   T.bd, p.eps, p.usequad, nx, ny, meta);
    M.nlast = total;
    BH_HIP(hipGetLastError());
    // append to the training set
    const int64_t old = M.nsamples, nn = old + total;
    DevBuf gx, gy;
    float *X = ensure<float>(gx, (size_t)nn * kMlpIn), *Y = ensure<float>(gy, (size_t)nn * kMlpOut);
    if (old) {
        BH_HIP(hipMemcpyAsync(X, M.sx.p, old * kMlpIn * sizeof(float), hipMemcpyDeviceToDevice, st));
        BH_HIP(hipMemcpyAsync(Y, M.sy.p, old * kMlpOut * sizeof(float), hipMemcpyDeviceToDevice, st));
    }
    BH_HIP(hipMemcpyAsync(X + old * kMlpIn, nx, (size_t)total * kMlpIn * sizeof(float), hipMemcpyDeviceToDevice, st));
    BH_HIP(hipMemcpyAsync(Y + old * kMlpOut, ny, (size_t)total * kMlpOut * sizeof(float), hipMemcpyDeviceToDevice,
                          st));
    BH_HIP(hipStreamSynchronize(st));
    if (M.sx.p) BH_HIP(hipFree(M.sx.p));
    if (M.sy.p) BH_HIP(hipFree(M.sy.p));
    M.sx = gx;
    M.sy = gy;
    M.nsamples = nn;
    return total;
}

void mlp_forward(bh_ctx *ctx, const float *d_x, int64_t n, float *d_y)
{
    Mlp &M = ctx->mlp;
    if (n <= 0) return;
    k_mlp_forward<<<(int)std::min<int64_t>((n + 63) / 64, (int64_t)ctx->ncu * 8), 64, 0, ctx->stream>>>(
        d_x, n, ptr<bf16>(M.wb), ptr<float>(M.bias), d_y);
    BH_HIP(hipGetLastError());
}

// Gradient of the mean squared error over the batch (x, y) at the current
// parameters: grad[kMlpParams], loss.
void mlp_grad(bh_ctx *ctx, const float *d_x, const float *d_y, int64_t n, float *d_grad, float *d_loss)
{
    Mlp &M = ctx->mlp;
    const int nw = (int)((n + 63) / 64);
    float *gpart = ensure<float>(M.gpart, (size_t)nw * kMlpParams);
    float *lpart = ensure<float>(M.lpart, nw);
    k_mlp_grad<<<nw, 64, 0, ctx->stream>>>(d_x, d_y, n, 1.f / (float)n, ptr<bf16>(M.wb), ptr<float>(M.bias), gpart,
                                           lpart);
    k_mlp_reduce<<<(kMlpParams + 255) / 256, 256, 0, ctx->stream>>>(gpart, nw, d_grad, lpart, d_loss);
    BH_HIP(hipGetLastError());
}

void mlp_gather(bh_ctx *ctx, uint64_t a, uint64_t c, int64_t first, int64_t cnt, float *bx, float *by)
{
    Mlp &M = ctx->mlp;
    k_gather_batch<<<(int)((cnt + 255) / 256), 256, 0, ctx->stream>>>(ptr<float>(M.sx), ptr<float>(M.sy), M.nsamples, a, c,
                                                                    first, cnt, 1.f / M.norm[0], 1.f / M.norm[3], bx,
                                                                    by);
    BH_HIP(hipGetLastError());
}

void mlp_adam(bh_ctx *ctx, const float *d_grad, float lr, int step)
{
    Mlp &M = ctx->mlp;
    float *m = ensure<float>(M.adam_m, kMlpParams), *v = ensure<float>(M.adam_v, kMlpParams);
    if (step == 1) {
        BH_HIP(hipMemsetAsync(m, 0, kMlpParams * sizeof(float), ctx->stream));
        BH_HIP(hipMemsetAsync(v, 0, kMlpParams * sizeof(float), ctx->stream));
    }
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    k_adam<<<(kMlpParams + 255) / 256, 256, 0, ctx->stream>>>(ptr<float>(M.params), d_grad, m, v, lr, b1, b2, 1e-8f, c1,
                                                              c2);
    mlp_pack(ctx);
    BH_HIP(hipGetLastError());
}

}  // namespace bh
