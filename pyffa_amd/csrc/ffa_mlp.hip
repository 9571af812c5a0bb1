// Field surrogate kernels (gfx950): training forward/backward, weight
// gradients, Adam + repacking, and batched evaluation.  The network and the
// matrix-core mapping are described in ffa_mlp.h.
//
// Training works on a batch of Bp points (n rounded up to FFA_MLP_KCHUNK;
// padding points get zero loss gradient):
//   k_mlp_fwd_bwd  one wave per 64 points: forward pass, squared error and
//                  its gradient, backward pass down to the first layer;
//                  writes the BF16 activations x_l and backward signals dz_l
//                  feature-major ([row][point]) and the FP32
//                  pre-activations in wave order (reread by the same wave);
//   k_mlp_wgrad    dW_l = x_l^T dz_l and db_l = sum dz_l per (layer, 32x32
//                  tile, chunk of FFA_MLP_KCHUNK points) on the matrix cores,
//                  one FP32 partial per chunk (no atomics);
//   k_mlp_update   fixed-order sum of the chunk partials, Adam, repack.
// Every sum has a fixed order, so a training run is bit-reproducible.
#include "ffa_mlp.h"

namespace {

using namespace mlp;

constexpr int XROWS = NFEAT + 4 * HID;        // x_0 .. x_4
constexpr int DROWS = 4 * HID + NOUT;         // dz_0 .. dz_3, dy
constexpr int NTILES = 8 + 3 * 16 + 4;        // weight-gradient tiles

__host__ __device__ constexpr int xoff(int l) { return l == 0 ? 0 : NFEAT + (l - 1) * HID; }
__host__ __device__ constexpr int doff(int l) { return l * HID; }

struct AdamF {
    float lr, b1, b2, eps, bc1, bc2;   // bc = 1 / (1 - beta^t)
};

__device__ __forceinline__ void pack_param(char *packed, int i, float w) {
    __bf16 *Wf = reinterpret_cast<__bf16 *>(packed);
    float *bias = reinterpret_cast<float *>(packed + FWD_BF16 * 2);
    __bf16 *Wb = reinterpret_cast<__bf16 *>(packed + FWD_BYTES);
    const __bf16 wb = (__bf16)w;
    // forward copy: W_0 and the hidden biases x KAPPA, W_4 x 1 / KAPPA
    if (i < PB0) {
        const int in = i / HID, out = i % HID;
        Wf[F0 + out * LD0 + in] = (__bf16)(w * KAPPA);
    } else if (i < PW1) {
        bias[i - PB0] = w * KAPPA;
    } else if (i < PB3 + HID && i >= PW1) {
        // layers 1..3: [W_l (HID*HID), b_l (HID)] blocks of HID*HID + HID
        const int l = 1 + (i - PW1) / (HID * HID + HID);
        const int j = (i - PW1) % (HID * HID + HID);
        if (j >= HID * HID) {
            bias[l * HID + (j - HID * HID)] = w * KAPPA;
        } else {
            const int in = j / HID, out = j % HID;
            const int fo = l == 1 ? F1 : (l == 2 ? F2 : F3);
            const int bo = l == 1 ? B1 : (l == 2 ? B2 : B3);
            Wf[fo + out * LDH + perm16(in)] = wb;
            Wb[bo + in * HID + perm16(out)] = wb;
        }
    } else if (i < PB4) {
        const int j = i - PW4, in = j / NOUT, out = j % NOUT;
        Wf[F4 + out * LDH + perm16(in)] = (__bf16)(w * INV_KAPPA);
        Wb[B4 + in * 16 + out] = wb;
    } else {
        bias[4 * HID + (i - PB4)] = w;
    }
}

// mode 0: pack params; 1: grad = sum of partials (or grad_in) -> grad_out;
// 2: as 1, then Adam on params/m/v and repack.
__global__ __launch_bounds__(256) void k_mlp_update(float *__restrict__ params, const float *__restrict__ partials,
                                                   int S, const float *__restrict__ grad_in,
                                                   float *__restrict__ grad_out, float *__restrict__ mom,
                                                   float *__restrict__ vel, char *__restrict__ packed, AdamF hp,
                                                   int mode, const float *__restrict__ lossw, int64_t nw,
                                                   float loss_scale, float *__restrict__ loss_out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (loss_out != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        float s = 0.0f;
        for (int64_t k = 0; k < nw; ++k) s += lossw[k];
        loss_out[0] = s * loss_scale;
    }
    if (i >= NPARAM) return;
    float g = 0.0f;
    if (mode >= 1) {
        if (grad_in != nullptr) {
            g = grad_in[i];
        } else {
            for (int s = 0; s < S; ++s) g += partials[(int64_t)s * NPARAM + i];
            // layers 1..4 took their inputs as the scaled activations KAPPA silu
            const bool scaled_in = (i >= PW1 && i < PB1) || (i >= PW2 && i < PB2) || (i >= PW3 && i < PB3) ||
                                   (i >= PW4 && i < PB4);
            if (scaled_in) g *= INV_KAPPA;
        }
        if (grad_out != nullptr) grad_out[i] = g;
        if (mode == 1) return;      // gradient only: params, moments and packed are not touched (may be null)
    }
    float w = params[i];
    if (mode == 2) {
        const float m = hp.b1 * mom[i] + (1.0f - hp.b1) * g;
        const float v = hp.b2 * vel[i] + (1.0f - hp.b2) * g * g;
        mom[i] = m;
        vel[i] = v;
        w -= hp.lr * (m * hp.bc1) / (sqrtf(v * hp.bc2) + hp.eps);
        params[i] = w;
    }
    pack_param(packed, i, w);
}

// dz_l = dx * silu'(z_l) for input block ib of the layer above: BF16 operands
// for the next backward product and the feature-major copy for dW.
__device__ __forceinline__ void dz_block(const f32x16 (&dx)[2], const float *zw, int l, int ib, int lane, int h,
                                         bf16x8 (&g)[2][8], __bf16 *dzfm, int64_t Bp, const int64_t (&pt)[2]) {
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
            const float zv = zw[((((l * 2 + t) * 4 + ib) * 16) + rr) * 64 + lane];
            const __bf16 v = (__bf16)(dx[t][rr] * dsilu_k(zv));
            g[t][2 * ib + (rr >> 3)][rr & 7] = v;
            dzfm[(int64_t)(doff(l) + ib * 32 + acc_row(rr, h)) * Bp + pt[t]] = v;
        }
}

__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const char *__restrict__ packed, fdesc d,
                                                    const float *__restrict__ pts, const float *__restrict__ tgt,
                                                    int64_t n, int64_t Bp, __bf16 *__restrict__ xfm,
                                                    __bf16 *__restrict__ dzfm, float *__restrict__ zs,
                                                    float *__restrict__ lossw, float dy_scale) {
    const __bf16 *Wf = reinterpret_cast<const __bf16 *>(packed);
    const float *bias = reinterpret_cast<const float *>(packed + FWD_BF16 * 2);
    const __bf16 *Wb = reinterpret_cast<const __bf16 *>(packed + FWD_BYTES);
    const int lane = threadIdx.x & 63, h = lane >> 5;
    int nl = lane & 31;
    const int64_t wave = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
    const int64_t p0 = wave * 64;
    const int64_t pown = p0 + lane;
    float r = d.r_c, fi = 0.0f, z = 0.0f;
    if (pown < n) {
        r = pts[pown];
        fi = pts[n + pown];
        z = pts[2 * n + pown];
    }
    float a, u, c;
    inputs(&d, r, fi, z, a, u, c);
    const float c2 = c * c;
    const float ax = __shfl_xor(a, 32), ux = __shfl_xor(u, 32), cx = __shfl_xor(c2, 32);
    const float a0 = h ? ax : a, a1 = h ? a : ax;
    const float u0 = h ? ux : u, u1 = h ? u : ux;
    const float q0 = h ? cx : c2, q1 = h ? c2 : cx;
    const int64_t pt[2] = {p0 + nl, p0 + 32 + nl};
    float *zw = zs + wave * (4 * 2 * 4 * 16 * 64);

    bf16x8 xa[2][8], xb[2][8];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
        xa[0][cc] = feature_chunk(&d, cc, h, a0, u0, q0);
        xa[1][cc] = feature_chunk(&d, cc, h, a1, u1, q1);
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int e = 0; e < 8; ++e) xfm[(int64_t)(16 * cc + 8 * h + e) * Bp + pt[t]] = xa[t][cc][e];
    }
    // forward, block by block; pre-activations to zs, activations to xfm
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        bf16x8 (&xi)[2][8] = (l & 1) ? xb : xa;
        bf16x8 (&xo)[2][8] = (l & 1) ? xa : xb;
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            f32x16 y[2];
            asm volatile("" : "+v"(nl));   // one block's weight loads at a time
            if (l == 0) dense_block<4, true>(Wf + F0, LD0, xi, ob, nl, h, y[0], y[1], bias);
            else dense_block<8, true>(Wf + (l == 1 ? F1 : (l == 2 ? F2 : F3)), LDH, xi, ob, nl, h, y[0], y[1],
                                      bias + l * HID);
            activate_block(y[0], y[1], ob, xo);
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int rr = 0; rr < 16; ++rr) {
                    zw[((((l * 2 + t) * 4 + ob) * 16) + rr) * 64 + lane] = y[t][rr];
                    xfm[(int64_t)(xoff(l + 1) + ob * 32 + acc_row(rr, h)) * Bp + pt[t]] = xo[t][2 * ob + (rr >> 3)][rr & 7];
                }
        }
    }
    // after four layers the activations x_4 are in xa
    f32x16 y[2];
    dense_block<8>(Wf + F4, LDH, xa, 0, nl, h, y[0], y[1]);
    // squared error and dy on lane half 0 (rows 0..2 = outputs)
    float sq = 0.0f;
    bf16x8 gy[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        gy[t] = bf16x8{};
        const bool valid = h == 0 && pt[t] < n;
#pragma unroll
        for (int k = 0; k < NOUT; ++k) {
            float e = 0.0f;
            if (valid) e = (y[t][k] + bias[4 * HID + k]) - tgt[(int64_t)k * n + pt[t]];
            sq += e * e;
            const __bf16 g = (__bf16)(e * dy_scale);
            if (h == 0) {
                gy[t][k] = g;
                dzfm[(int64_t)(doff(4) + k) * Bp + pt[t]] = g;
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) sq += __shfl_xor(sq, off);
    if (lane == 0) lossw[wave] = sq;
    // backward, block by block over the input features of each layer:
    // dx_4 = W4 dy; dz_l = dx_(l+1) * silu'(z_l); dx_l = W_l dz_l
    bf16x8 (&g4)[2][8] = xb;       // dz_3 operands
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
        const bf16x8 A = load8(Wb + B4 + (ib * 32 + nl) * 16 + h * 8);
        f32x16 dx[2] = {mfma(A, gy[0], f32x16{}), mfma(A, gy[1], f32x16{})};
        dz_block(dx, zw, 3, ib, lane, h, g4, dzfm, Bp, pt);
    }
#pragma unroll
    for (int l = 3; l >= 1; --l) {
        bf16x8 (&gi)[2][8] = (l & 1) ? xb : xa;    // dz_l
        bf16x8 (&go)[2][8] = (l & 1) ? xa : xb;    // dz_(l-1)
#pragma unroll
        for (int ib = 0; ib < 4; ++ib) {
            f32x16 dx[2];
            asm volatile("" : "+v"(nl));
            dense_block<8>(Wb + (l == 3 ? B3 : (l == 2 ? B2 : B1)), HID, gi, ib, nl, h, dx[0], dx[1]);
            dz_block(dx, zw, l - 1, ib, lane, h, go, dzfm, Bp, pt);
        }
    }
}

__global__ __launch_bounds__(64) void k_mlp_wgrad(const __bf16 *__restrict__ xfm, const __bf16 *__restrict__ dzfm,
                                                 int64_t Bp, float *__restrict__ partials) {
    const int tile = blockIdx.x, s = blockIdx.y;
    const int lane = threadIdx.x, nl = lane & 31, h = lane >> 5;
    int L, mb, nb;
    if (tile < 8) {
        L = 0; mb = tile >> 2; nb = tile & 3;
    } else if (tile < 56) {
        L = 1 + (tile - 8) / 16;
        mb = ((tile - 8) % 16) >> 2;
        nb = (tile - 8) & 3;
    } else {
        L = 4; mb = tile - 56; nb = 0;
    }
    const int col = nb * 32 + nl;
    const bool vcol = L < 4 || col < NOUT;
    const __bf16 *xa = xfm + (int64_t)(xoff(L) + mb * 32 + nl) * Bp + 8 * h;
    const __bf16 *db = dzfm + (int64_t)(doff(L) + (vcol ? col : 0)) * Bp + 8 * h;
    f32x16 acc = {};
    float bsum = 0.0f;
    const int64_t pbeg = (int64_t)s * FFA_MLP_KCHUNK;
    for (int64_t p = pbeg; p < pbeg + FFA_MLP_KCHUNK; p += 16) {
        const bf16x8 A = load8(xa + p);
        bf16x8 B = bf16x8{};
        if (vcol) B = load8(db + p);
        acc = mfma(A, B, acc);
        if (mb == 0) {
#pragma unroll
            for (int e = 0; e < 8; ++e) bsum += (float)B[e];
        }
    }
    const int outdim = L < 4 ? HID : NOUT;
    const int pw = L == 0 ? PW0 : (L == 1 ? PW1 : (L == 2 ? PW2 : (L == 3 ? PW3 : PW4)));
    const int pb = L == 0 ? PB0 : (L == 1 ? PB1 : (L == 2 ? PB2 : (L == 3 ? PB3 : PB4)));
    float *out = partials + (int64_t)s * NPARAM;
    if (vcol) {
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) out[pw + (mb * 32 + acc_row(rr, h)) * outdim + col] = acc[rr];
    }
    if (mb == 0) {
        bsum += __shfl_xor(bsum, 32);
        if (h == 0 && vcol) out[pb + col] = bsum;
    }
}

constexpr int EVAL_BLOCK = 512;

__global__ __launch_bounds__(EVAL_BLOCK) void k_mlp_eval(const char *__restrict__ packed, fdesc d,
                                                        const float *__restrict__ pts, int64_t n,
                                                        float *__restrict__ out, int mode) {
    __shared__ __attribute__((aligned(16))) char lds[FWD_BYTES];
    for (int k = threadIdx.x; k < FWD_BYTES / 16; k += EVAL_BLOCK)
        reinterpret_cast<uint4 *>(lds)[k] = reinterpret_cast<const uint4 *>(packed)[k];
    __syncthreads();
    const __bf16 *W = reinterpret_cast<const __bf16 *>(lds);
    const float *b = reinterpret_cast<const float *>(lds + FWD_BF16 * 2);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t stride = (int64_t)gridDim.x * (EVAL_BLOCK / 64) * 64;
    for (int64_t base = ((int64_t)blockIdx.x * (EVAL_BLOCK / 64) + wv) * 64; base < n; base += stride) {
        const int64_t p = base + lane;
        const bool valid = p < n;
        float r = d.r_c, fi = 0.0f, z = 0.0f;
        if (valid) {
            r = pts[p];
            fi = pts[n + p];
            z = pts[2 * n + p];
        }
        float a, u, c;
        inputs(&d, r, fi, z, a, u, c);
        float o[3];
        forward_wave(W, b, &d, a, u, c * c, o);
        if (valid) {
            if (mode == 0) {
                float bz, bfi, br;
                field_from_outputs(&d, c, o, bz, bfi, br);
                out[p] = bz;
                out[n + p] = bfi;
                out[2 * n + p] = br;
            } else {
                out[p] = o[0];
                out[n + p] = o[1];
                out[2 * n + p] = o[2];
            }
        }
    }
}

}  // namespace

namespace mlp {

fdesc fdesc_from(const ffa_mlp_desc &d) {
    fdesc f{};
    f.r_c = (float)d.r_c;
    f.inv_rh = (float)d.inv_rh;
    f.inv_sector = (float)d.inv_sector;
    f.zt = (float)d.zt;
    f.inv_zt = (float)(1.0 / d.zt);
    for (int k = 0; k < 3; ++k) f.S[k] = (float)d.out_scale[k];
    f.delta = (float)d.out_delta;
    for (int k = 0; k < NFREQ; ++k) {
        f.m[k] = d.m[k];
        f.wa[k] = d.wa[k];
        f.wc[k] = d.wc[k];
    }
    f.fallback = d.fallback;
    return f;
}

size_t workspace_layout(int64_t n, ffa_mlp_layout &L) {
    const int64_t Bp = (n + FFA_MLP_KCHUNK - 1) / FFA_MLP_KCHUNK * FFA_MLP_KCHUNK;
    const int64_t S = Bp / FFA_MLP_KCHUNK, nw = Bp / 64;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t o = 0;
    L.batch_padded = Bp;
    L.xfm = o; o = al(o + (size_t)XROWS * Bp * 2);
    L.dzfm = o; o = al(o + (size_t)DROWS * Bp * 2);
    L.zs = o; o = al(o + (size_t)nw * 4 * 2 * 4 * 16 * 64 * 4);
    L.partials = o; o = al(o + (size_t)S * NPARAM * 4);
    L.loss = o; o = al(o + (size_t)nw * 4);
    L.total = o;
    return o;
}

}  // namespace mlp

hipError_t mlp_upload_desc(ffa_ctx_s *c, const ffa_mlp_desc *d) {
    const fdesc fd = mlp::fdesc_from(*d);
    static_assert(sizeof(fdesc) <= FFA_CTX_MLP_BYTES, "surrogate descriptor block overflows");
    hipError_t e = hipMemcpyAsync(c->dev + FFA_CTX_MLP_OFF, &fd, sizeof fd, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

hipError_t launch_mlp_pack(ffa_ctx_s *c, const float *params, void *packed) {
    hipLaunchKernelGGL(k_mlp_update, dim3((NPARAM + 255) / 256), dim3(256), 0, c->stream, const_cast<float *>(params),
                       nullptr, 0, nullptr, nullptr, nullptr, nullptr, static_cast<char *>(packed), AdamF{}, 0,
                       nullptr, (int64_t)0, 0.0f, nullptr);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(ffa_ctx_s *c, const ffa_mlp_desc *d, const void *packed, const float *pts,
                           const float *tgt, int64_t n, void *ws, float *params, float *grad, float *mom, float *vel,
                           void *packed_out, const ffa_adam *hp, float *loss) {
    ffa_mlp_layout L;
    mlp::workspace_layout(n, L);
    char *w = static_cast<char *>(ws);
    const int64_t Bp = L.batch_padded, S = Bp / FFA_MLP_KCHUNK, nw = Bp / 64;
    const fdesc fd = mlp::fdesc_from(*d);
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3((unsigned)(Bp / 256)), dim3(256), 0, c->stream,
                       static_cast<const char *>(packed), fd, pts, tgt, n, Bp, reinterpret_cast<__bf16 *>(w + L.xfm),
                       reinterpret_cast<__bf16 *>(w + L.dzfm), reinterpret_cast<float *>(w + L.zs),
                       reinterpret_cast<float *>(w + L.loss), (float)(2.0 / (3.0 * (double)n)));
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(NTILES, (unsigned)S), dim3(64), 0, c->stream,
                       reinterpret_cast<const __bf16 *>(w + L.xfm), reinterpret_cast<const __bf16 *>(w + L.dzfm), Bp,
                       reinterpret_cast<float *>(w + L.partials));
    AdamF a{};
    int mode = 1;
    if (hp != nullptr) {
        mode = 2;
        a.lr = hp->lr;
        a.b1 = hp->beta1;
        a.b2 = hp->beta2;
        a.eps = hp->eps;
        a.bc1 = (float)(1.0 / (1.0 - pow((double)hp->beta1, (double)hp->step)));
        a.bc2 = (float)(1.0 / (1.0 - pow((double)hp->beta2, (double)hp->step)));
    }
    hipLaunchKernelGGL(k_mlp_update, dim3((NPARAM + 255) / 256), dim3(256), 0, c->stream, params,
                       reinterpret_cast<const float *>(w + L.partials), (int)S, nullptr, grad, mom, vel,
                       static_cast<char *>(packed_out), a, mode, reinterpret_cast<const float *>(w + L.loss), nw,
                       (float)(1.0 / (3.0 * (double)n)), loss);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(ffa_ctx_s *c, float *params, const float *grad, float *mom, float *vel, void *packed,
                           const ffa_adam *hp) {
    AdamF a{hp->lr, hp->beta1, hp->beta2, hp->eps,
            (float)(1.0 / (1.0 - pow((double)hp->beta1, (double)hp->step))),
            (float)(1.0 / (1.0 - pow((double)hp->beta2, (double)hp->step)))};
    hipLaunchKernelGGL(k_mlp_update, dim3((NPARAM + 255) / 256), dim3(256), 0, c->stream, params, nullptr, 0, grad,
                       nullptr, mom, vel, static_cast<char *>(packed), a, 2, nullptr, (int64_t)0, 0.0f, nullptr);
    return hipGetLastError();
}

hipError_t launch_mlp_eval(ffa_ctx_s *c, const ffa_mlp_desc *d, const void *packed, const float *pts, int64_t n,
                           float *out, int mode) {
    if (n == 0) return hipSuccess;
    const fdesc fd = mlp::fdesc_from(*d);
    // the weights take 133 KB of LDS, so one workgroup fits a CU: launch one
    // per CU and let each loop over its batches, loading the weights once
    int ncu = 0;
    hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device);
    if (e != hipSuccess) return e;
    int64_t g = (n + EVAL_BLOCK - 1) / EVAL_BLOCK;
    if (g > ncu) g = ncu > 0 ? ncu : 1;
    hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)g), dim3(EVAL_BLOCK), 0, c->stream,
                       static_cast<const char *>(packed), fd, pts, n, out, mode);
    return hipGetLastError();
}
