// Field surrogate on the gfx950 matrix cores: layouts and the wave-level
// forward pass shared by the push kernel and the surrogate kernels.
//
// Matrix-core mapping.  A layer y = x W + b over the 64 points of a wave is
// computed transposed, y^T = W^T x^T, with v_mfma_f32_32x32x16_bf16: the A
// operand is a 32(out) x 16(in) block of W^T, the B operand a 16(in) x 32
// (points) block of x^T, so the points of the wave form two 32-column tiles
// and each A fragment is used for both.  Lane l of the accumulator holds
// point column l % 32 and the 16 rows 8 i + 4 (l / 32) + j (register
// 4 i + j).  Those 16 values become, after bias + SiLU + BF16 rounding, the
// lane's B operands of the next layer for two 16-row K chunks without any
// data exchange between lanes: element e = 4 a + j of K chunk c in lane half
// h then stands for input feature 16 c + 8 a + 4 h + j.  The packed W^T of
// layers 1..4 stores that feature at column 16 c + 8 h + 4 a + j (bits 2 and
// 3 of the in-chunk index swapped, perm16), so an A fragment is one
// contiguous 16-byte load.  The backward pass uses the same mapping with W
// (in x out) in place of W^T.
//
// Scaled pre-activations.  The packed forward weights carry the factor
// KAPPA = -log2(e): W_0 and the hidden biases are multiplied by it and W_4 by
// 1 / KAPPA, so the matrix cores return y' = KAPPA y for every hidden unit.
// Then silu(y) = y / (1 + e^-y) is y' / (1 + 2^y') / KAPPA: the activation
// a' = y' / (1 + 2^y') = KAPPA silu(y) costs an exp2, an add, a reciprocal
// and one multiply, and the 1 / KAPPA it owes is already in the next layer
// (hidden W_l are packed unscaled: KAPPA y_(l+1) = W_(l+1) a' + KAPPA b).
// The backward weights W (in x out) are packed unscaled; the weight
// gradients taken from the scaled activations are divided by KAPPA when the
// chunks are summed.
#pragma once
#include "ffa_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace mlp {

constexpr int NFREQ = 32, NFEAT = 64, HID = 128, NOUT = 3;
// forward packed layout: BF16 W^T rows (padded by 8 elements so the 32 rows
// read by one fragment load start on different LDS banks), then FP32 biases
constexpr int LD0 = NFEAT + 8, LDH = HID + 8;
constexpr int F0 = 0;
constexpr int F1 = F0 + HID * LD0;
constexpr int F2 = F1 + HID * LDH;
constexpr int F3 = F2 + HID * LDH;
constexpr int F4 = F3 + HID * LDH;
constexpr int FWD_BF16 = F4 + 32 * LDH;          // W4^T padded to 32 rows
constexpr int NBIAS = 4 * HID + 32;
constexpr int FWD_BYTES = FWD_BF16 * 2 + NBIAS * 4;
// backward packed layout: BF16 W (in x out-permuted), layers 4, 3, 2, 1
constexpr int B4 = 0;                            // 128 x 16 (3 valid columns)
constexpr int B3 = B4 + HID * 16;
constexpr int B2 = B3 + HID * HID;
constexpr int B1 = B2 + HID * HID;
constexpr int BWD_BF16 = B1 + HID * HID;
constexpr int PACKED_BYTES = FWD_BYTES + BWD_BF16 * 2;
// FP32 master parameters
constexpr int PW0 = 0, PB0 = PW0 + NFEAT * HID;
constexpr int PW1 = PB0 + HID, PB1 = PW1 + HID * HID;
constexpr int PW2 = PB1 + HID, PB2 = PW2 + HID * HID;
constexpr int PW3 = PB2 + HID, PB3 = PW3 + HID * HID;
constexpr int PW4 = PB3 + HID, PB4 = PW4 + HID * NOUT;
constexpr int NPARAM = PB4 + NOUT;
static_assert(NPARAM == FFA_MLP_NPARAM, "parameter count");
static_assert(PACKED_BYTES == FFA_MLP_PACKED_BYTES, "packed size");
static_assert(FWD_BYTES % 16 == 0, "LDS copy in 16-byte units");

__host__ __device__ constexpr int perm16(int k) { return (k & ~12) | ((k & 4) << 1) | ((k & 8) >> 1); }

// FP32 copy of ffa_mlp_desc used on the device
struct fdesc {
    float r_c, inv_rh, inv_sector, zt, inv_zt;
    float S[3];
    float delta;                   // output map: 0 -> o, > 0 -> delta sinh(o)
    float m[NFREQ], wa[NFREQ], wc[NFREQ];
    int32_t fallback, pad[3];
};

constexpr float KAPPA = -1.4426950408889634f;          // -log2(e)
constexpr float INV_KAPPA = -0.6931471805599453f;      // -ln(2)

// KAPPA silu(y) of a scaled pre-activation yk = KAPPA y (2^yk = e^-y)
__device__ __forceinline__ float silu_k(float yk) {
    return yk * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(yk));
}

// silu'(y) = sigma (1 + y (1 - sigma)) of a scaled pre-activation yk = KAPPA y
__device__ __forceinline__ float dsilu_k(float yk) {
    const float s = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(yk));
    const float y = yk * INV_KAPPA;
    return s * (1.0f + y * (1.0f - s));
}

// Taylor-part field of a point from the network outputs o at c = z / zt
__device__ __forceinline__ void field_from_outputs(const fdesc *d, float c, const float o[3], float &Bz, float &Bfi,
                                                   float &Br) {
    float g0 = o[0], g1 = o[1], g2 = o[2];
    if (d->delta > 0.0f) {
        g0 = 0.5f * d->delta * (__expf(g0) - __expf(-g0));
        g1 = 0.5f * d->delta * (__expf(g1) - __expf(-g1));
        g2 = 0.5f * d->delta * (__expf(g2) - __expf(-g2));
    }
    const float c2 = c * c, c3 = c2 * c;
    Bz = d->S[0] * c2 * g0;
    Bfi = d->S[1] * c3 * g1;
    Br = d->S[2] * c3 * g2;
}

// network inputs of a point
__device__ __forceinline__ void inputs(const fdesc *d, float r, float fi, float z, float &a, float &u, float &c) {
    a = (r - d->r_c) * d->inv_rh;
    const float v = fi * d->inv_sector;
    u = v - floorf(v);
    c = z * d->inv_zt;
}

// Fourier features of one 16-row K chunk for lane half h: elements 2q, 2q+1
// are sin, cos of frequency 8 c + 4 h + q.
__device__ __forceinline__ bf16x8 feature_chunk(const fdesc *d, int c, int h, float a, float u, float c2) {
    bf16x8 x;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        // constant table indices (uniform loads), lane half selects
        const int f0 = 8 * c + q, f1 = f0 + 4;
        const float mf = h ? d->m[f1] : d->m[f0];
        const float waf = h ? d->wa[f1] : d->wa[f0];
        const float wcf = h ? d->wc[f1] : d->wc[f0];
        float rev = mf * u + waf * a + wcf * c2;
        rev -= floorf(rev);
        x[2 * q] = (__bf16)__builtin_amdgcn_sinf(rev);
        x[2 * q + 1] = (__bf16)__builtin_amdgcn_cosf(rev);
    }
    return x;
}

__device__ __forceinline__ bf16x8 load8(const __bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// One 32-row output block ob of a layer for both point tiles:
// y_t = sum over NCH K chunks of W^T(ob rows, chunk c) x^T(chunk c, tile t).
// Layers are evaluated block by block so that only two accumulators are
// live next to the input and output operands (no spills at 256 registers).
// row of accumulator register r (0..15) in lane half h
__device__ __forceinline__ int acc_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

// BIAS: the accumulators start from the block's biases b (rows ob * 32 ..),
// so the layer's pre-activations leave the matrix cores complete.
template <int NCH, bool BIAS = false>
__device__ __forceinline__ void dense_block(const __bf16 *Wl, int ld, const bf16x8 (&x)[2][8], int ob, int n, int h,
                                            f32x16 &y0, f32x16 &y1, const float *b = nullptr) {
    if constexpr (BIAS) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            y0[r] = b[ob * 32 + acc_row(r, h)];
            y1[r] = y0[r];
        }
    } else {
        y0 = f32x16{};
        y1 = f32x16{};
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const bf16x8 A = load8(Wl + (ob * 32 + n) * ld + c * 16 + h * 8);
        y0 = mfma(A, x[0][c], y0);
        y1 = mfma(A, x[1][c], y1);
    }
}

// SiLU + BF16 of the scaled pre-activations (bias included) of output block
// ob -> B operands (K chunks 2 ob, 2 ob + 1) of the next layer.
__device__ __forceinline__ void activate_block(const f32x16 &y0, const f32x16 &y1, int ob, bf16x8 (&x)[2][8]) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        x[0][2 * ob + (r >> 3)][r & 7] = (__bf16)silu_k(y0[r]);
        x[1][2 * ob + (r >> 3)][r & 7] = (__bf16)silu_k(y1[r]);
    }
}

// Software-pipelined over the output blocks: the matrix-core products of
// block ob + 1 are issued before the SiLU of block ob, so the VALU work of
// one block runs while the matrix cores compute the next (the two only share
// the layer's input operands).
template <int NCH>
__device__ __forceinline__ void dense_silu(const __bf16 *Wl, int ld, const float *b, const bf16x8 (&xin)[2][8],
                                           bf16x8 (&xout)[2][8], int n, int h) {
    f32x16 y[2][2];
    dense_block<NCH, true>(Wl, ld, xin, 0, n, h, y[0][0], y[0][1], b);
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        if (ob + 1 < 4) dense_block<NCH, true>(Wl, ld, xin, ob + 1, n, h, y[(ob + 1) & 1][0], y[(ob + 1) & 1][1], b);
        activate_block(y[ob & 1][0], y[ob & 1][1], ob, xout);
    }
}

// Network outputs o[0..2] of this lane's point.  Every lane of the wave must
// call it (the 64 points of the wave share the matrix-core work).  W: packed
// forward weights (LDS or global), b: the FP32 biases behind them.
__device__ __forceinline__ void forward_wave(const __bf16 *W, const float *b, const fdesc *d, float a, float u,
                                             float c2, float o[3]) {
    const int lane = threadIdx.x & 63, h = lane >> 5;
    int n = lane & 31;
    // Opaque per call: keeps the 120 weight-fragment loads inside the
    // caller's loop instead of being hoisted (and spilled) as invariants.
    asm volatile("" : "+v"(n));
    const float ax = __shfl_xor(a, 32), ux = __shfl_xor(u, 32), cx = __shfl_xor(c2, 32);
    // tile t holds points 32 t + n; lane half h owns the point of tile h
    const float a0 = h ? ax : a, a1 = h ? a : ax;
    const float u0 = h ? ux : u, u1 = h ? u : ux;
    const float q0 = h ? cx : c2, q1 = h ? c2 : cx;
    bf16x8 xa[2][8], xb[2][8];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        xa[0][c] = feature_chunk(d, c, h, a0, u0, q0);
        xa[1][c] = feature_chunk(d, c, h, a1, u1, q1);
    }
    dense_silu<4>(W + F0, LD0, b, xa, xb, n, h);
    dense_silu<8>(W + F1, LDH, b + HID, xb, xa, n, h);
    dense_silu<8>(W + F2, LDH, b + 2 * HID, xa, xb, n, h);
    dense_silu<8>(W + F3, LDH, b + 3 * HID, xb, xa, n, h);
    f32x16 y0, y1;
    dense_block<8>(W + F4, LDH, xa, 0, n, h, y0, y1);   // W4^T: one 32-row block
    // outputs 0..2 are rows 0..2: registers 0..2 of lane half 0
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float v0 = y0[k] + b[4 * HID + k], v1 = y1[k] + b[4 * HID + k];
        const float w = __shfl_xor(v1, 32);
        o[k] = h ? w : v0;
    }
}

}  // namespace mlp
