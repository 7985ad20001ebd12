// guess.hip -- learned initial guess for the Krylov solve (no counterpart in
// the reference; it replaces only the starting vector, which the reference
// takes as the previous step's Z_iF, Solver.hpp:1247-1255).
//
// A per-DOF multilayer perceptron 16 -> 64 -> 64 -> 1 (ReLU) predicts the
// increment of the face traces over one time step from 2-hop features of the
// warm-start residual r0 = b - A x_prev and of the previous increments:
//   f0..2 = r0, A r0, A^2 r0 (scaled by s_r = RMS(r0))
//   f3..5 = dx1, dx2, A dx1   (scaled by s_x = RMS(dx1))
//   f6 = Dirichlet row, f7 = component index, f8 = log10(s_r / s_x) / 4
// The prediction d = s_x * y is a direction: the starting vector is the
// combination x_prev + a_0 d + sum_j a_j dx_j over the network's direction and
// the last `history_depth` increments dx_1 (newest) .. dx_h, of least residual
// norm (normal equations solved by pivoted Cholesky with truncation), used
// only if it lowers the initial residual; the Krylov solve then judges
// convergence as before.
//
// Arithmetic: BF16 v_mfma_f32_16x16x16bf16_1k with FP32 accumulation.  Every
// layer is evaluated transposed, H^T = W^T X^T, so the D operand of one MFMA
// (lane l holds rows 4(l/16)+t, column l%16) is already the B operand of the
// next (lane l holds k = 4(l/16)+t, n = l%16): a wavefront carries a tile of
// 16 samples through the network in registers.  The weight gradients
// (reductions over samples) transpose their operands through LDS.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace mhfem {

constexpr int GB = 256;  // work-group: 4 wavefronts
constexpr int NF = 16, NH = 64;
// parameter vector layout (FP32 master copy and gradients)
constexpr int P_W1 = 0;                 // [NF][NH]
constexpr int P_B1 = P_W1 + NF * NH;    // [NH]
constexpr int P_W2 = P_B1 + NH;         // [NH][NH]
constexpr int P_B2 = P_W2 + NH * NH;    // [NH]
constexpr int P_W3 = P_B2 + NH;         // [NH]
constexpr int P_B3 = P_W3 + NH;         // [1]
constexpr int NPARAM = P_B3 + 1;        // 5313
constexpr int NPARAM_PAD = 5376;      // a multiple of 64 (k_adam takes 64 parameters per work-group)

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (NaN kept quiet)
__host__ __device__ inline unsigned short f2bf(float f)
{
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short h)
{
    unsigned u = (unsigned)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ float4_t mfma(short4_t a, short4_t b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// BF16 copies of the weights in the operand layouts the kernels load:
//   A1[mb]      lane l: W1[4q+t][16mb+i]          (i = l%16, q = l/16)
//   A2[mb][kb]  lane l: W2[16kb+4q+t][16mb+i]
//   A3[kb]      lane l: W3[16kb+4q+t] if i == 0 else 0
//   A3b[mb]     lane l: W3[16mb+i]    if q == 0 && t == 0 else 0   (backward)
//   A2b[mb][kb] lane l: W2[16mb+i][16kb+4q+t]                       (backward)
struct PackedWeights {
    short4_t A1[4][64];
    short4_t A2[16][64];
    short4_t A3[4][64];
    short4_t A3b[4][64];
    short4_t A2b[16][64];
    float b1[NH], b2[NH], b3;
};

__global__ void __launch_bounds__(64) k_pack_weights(const float *__restrict__ P, PackedWeights *__restrict__ W)
{
    const int l = threadIdx.x, i = l & 15, q = l >> 4;
    for (int mb = 0; mb < 4; mb++) {
        short4_t a;
        for (int t = 0; t < 4; t++) a[t] = (short)f2bf(P[P_W1 + (4 * q + t) * NH + 16 * mb + i]);
        W->A1[mb][l] = a;
        short4_t a3, a3b;
        for (int t = 0; t < 4; t++) {
            a3[t] = (short)(i == 0 ? f2bf(P[P_W3 + 16 * mb + 4 * q + t]) : 0);
            a3b[t] = (short)((q == 0 && t == 0) ? f2bf(P[P_W3 + 16 * mb + i]) : 0);
        }
        W->A3[mb][l] = a3;
        W->A3b[mb][l] = a3b;
        for (int kb = 0; kb < 4; kb++) {
            short4_t a2, a2b;
            for (int t = 0; t < 4; t++) {
                a2[t] = (short)f2bf(P[P_W2 + (16 * kb + 4 * q + t) * NH + 16 * mb + i]);
                a2b[t] = (short)f2bf(P[P_W2 + (16 * mb + i) * NH + 16 * kb + 4 * q + t]);
            }
            W->A2[mb * 4 + kb][l] = a2;
            W->A2b[mb * 4 + kb][l] = a2b;
        }
    }
    W->b1[l] = P[P_B1 + l];
    W->b2[l] = P[P_B2 + l];
    if (l == 0) W->b3 = P[P_B3];
}

// forward pass of one 16-sample tile; the activations stay in registers
struct Tile {
    short4_t X;       // B operand of layer 1
    short4_t H1[4];   // relu(z1) as bf16, B operand of layer 2
    short4_t H2[4];   // relu(z2) as bf16, B operand of layer 3
    float out;        // lanes 0..15: prediction of sample l
};

__device__ __forceinline__ short4_t relu_bias_bf16(float4_t z, const float *b, int base)
{
    short4_t h;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const float v = z[t] + b[base + t];
        h[t] = (short)f2bf(v > 0.f ? v : 0.f);
    }
    return h;
}

__device__ __forceinline__ void tile_forward(const PackedWeights &W, const short4_t (&A1)[4],
                                             const short4_t (&A2)[16], const short4_t (&A3)[4], Tile &T)
{
    const int l = threadIdx.x & 63, q = l >> 4;
    const float4_t zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mb = 0; mb < 4; mb++) T.H1[mb] = relu_bias_bf16(mfma(A1[mb], T.X, zero), W.b1, 16 * mb + 4 * q);
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        float4_t acc = zero;
#pragma unroll
        for (int kb = 0; kb < 4; kb++) acc = mfma(A2[mb * 4 + kb], T.H1[kb], acc);
        T.H2[mb] = relu_bias_bf16(acc, W.b2, 16 * mb + 4 * q);
    }
    float4_t o = zero;
#pragma unroll
    for (int kb = 0; kb < 4; kb++) o = mfma(A3[kb], T.H2[kb], o);
    T.out = o[0] + W.b3;  // row 0 of the output block: valid in lanes 0..15
}

__device__ __forceinline__ short4_t load_features(const unsigned short *__restrict__ X, long rows, long s0)
{
    const int l = threadIdx.x & 63, i = l & 15, q = l >> 4;
    const long s = s0 + i;
    short4_t x = {0, 0, 0, 0};
    if (s < rows) {
        const short4_t *p = reinterpret_cast<const short4_t *>(X + s * NF + 4 * q);
        x = *p;
    }
    return x;
}

// y[s] = MLP(X[s]) for all samples
__global__ void __launch_bounds__(GB) k_mlp_forward(long rows, const unsigned short *__restrict__ X,
                                                    const PackedWeights *__restrict__ Wg, float *__restrict__ y)
{
    const int l = threadIdx.x & 63;
    const PackedWeights &W = *Wg;
    short4_t A1[4], A2[16], A3[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        A1[k] = W.A1[k][l];
        A3[k] = W.A3[k][l];
    }
#pragma unroll
    for (int k = 0; k < 16; k++) A2[k] = W.A2[k][l];
    const long ntiles = (rows + 15) / 16;
    const long wave = (long)blockIdx.x * (GB / 64) + (threadIdx.x >> 6);
    const long nwaves = (long)gridDim.x * (GB / 64);
    for (long tile = wave; tile < ntiles; tile += nwaves) {
        Tile T;
        T.X = load_features(X, rows, tile * 16);
        tile_forward(W, A1, A2, A3, T);
        if (l < 16 && tile * 16 + l < rows) y[tile * 16 + l] = T.out;
    }
}

// Gradient of L = (1/rows) sum_s (y_s - t_s)^2 with respect to all
// parameters; one partial vector per work-group, plus the partial loss.
// The transposes of k_mlp_grad exchange values between the lanes of one
// wavefront through that wavefront's own LDS buffer.  The LDS executes a
// wavefront's operations in issue order, so no work-group barrier is needed:
// a wavefront-scope fence keeps the compiler from moving the accesses.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__global__ void __launch_bounds__(GB) k_mlp_grad(long rows, const unsigned short *__restrict__ X,
                                                 const float *__restrict__ target,
                                                 const PackedWeights *__restrict__ Wg, float inv_rows,
                                                 float *__restrict__ partial, long tile_stride, long tile_offset)
{
    const int l = threadIdx.x & 63, i = l & 15, q = l >> 4, wid = threadIdx.x >> 6;
    const PackedWeights &W = *Wg;
    __shared__ float tr[GB / 64][2][16][17];  // per-wave transpose buffers (padded)
    short4_t A1[4], A2[16], A3[4], A3b[4], A2b[16];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        A1[k] = W.A1[k][l];
        A3[k] = W.A3[k][l];
        A3b[k] = W.A3b[k][l];
    }
#pragma unroll
    for (int k = 0; k < 16; k++) {
        A2[k] = W.A2[k][l];
        A2b[k] = W.A2b[k][l];
    }
    const float4_t zero = {0.f, 0.f, 0.f, 0.f};
    float4_t gW2[16], gW1[4];
#pragma unroll
    for (int k = 0; k < 16; k++) gW2[k] = zero;
#pragma unroll
    for (int k = 0; k < 4; k++) gW1[k] = zero;
    float gW3[4][4] = {}, gb2[4][4] = {}, gb1[4][4] = {};
    float gb3 = 0.f, loss = 0.f;

    // every wavefront runs the same number of tiles (tiles past the end are
    // all-zero and contribute nothing)
    // the epoch uses every tile_stride-th 16-sample tile, starting at tile_offset
    const long ntiles = (rows + 15) / 16;
    const long nsub = tile_offset < ntiles ? (ntiles - tile_offset + tile_stride - 1) / tile_stride : 0;
    const long wave = (long)blockIdx.x * (GB / 64) + wid;
    const long nwaves = (long)gridDim.x * (GB / 64);
    const long iters = (nsub + nwaves - 1) / nwaves;
    // The inputs of a tile (features in both operand layouts, targets) are
    // loaded one tile ahead, so their latency hides behind the previous
    // tile's arithmetic.
    auto tile_start = [&](long it) {
        const long tsub = wave + it * nwaves;
        return tsub < nsub ? (tsub * tile_stride + tile_offset) * 16 : rows;
    };
    auto load_ax = [&](long s0) {
        short4_t ax;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const long s = s0 + 4 * q + t;
            ax[t] = (short)(s < rows ? X[s * NF + i] : 0);
        }
        return ax;
    };
    long s_next = tile_start(0);
    short4_t X_next = load_features(X, rows, s_next), AX_next = load_ax(s_next);
    float t_next = (l < 16 && s_next + l < rows) ? target[s_next + l] : 0.f;
    for (long it = 0; it < iters; it++) {
        const long s0 = s_next;
        Tile T;
        T.X = X_next;
        const short4_t AX = AX_next;
        const float tgt = t_next;
        if (it + 1 < iters) {
            s_next = tile_start(it + 1);
            X_next = load_features(X, rows, s_next);
            AX_next = load_ax(s_next);
            t_next = (l < 16 && s_next + l < rows) ? target[s_next + l] : 0.f;
        }
        tile_forward(W, A1, A2, A3, T);
        // dL/dy for the sample of lanes 0..15, broadcast to lanes with the same i
        float g = 0.f;
        if (l < 16 && s0 + l < rows) {
            const float d = T.out - tgt;
            loss += d * d;
            g = 2.f * d * inv_rows;
        }
        gb3 += g;
        const float gi = __shfl(g, i, 64);
        short4_t Bg = {0, 0, 0, 0};
        if (q == 0) Bg[0] = (short)f2bf(g);
        // layer 3 backward: dW3 (matrix-vector, per lane), g_z2
        short4_t gZ2[4];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const float4_t gh = mfma(A3b[mb], Bg, zero);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const float h2 = bf2f((unsigned short)T.H2[mb][t]);
                gW3[mb][t] += h2 * gi;
                const float gz = h2 > 0.f ? gh[t] : 0.f;
                gb2[mb][t] += gz;
                gZ2[mb][t] = (short)f2bf(gz);
            }
        }
        // layer 2 backward: g_z1
        short4_t gZ1[4];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            float4_t acc = zero;
#pragma unroll
            for (int kb = 0; kb < 4; kb++) acc = mfma(A2b[mb * 4 + kb], gZ2[kb], acc);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const float h1 = bf2f((unsigned short)T.H1[mb][t]);
                const float gz = h1 > 0.f ? acc[t] : 0.f;
                gb1[mb][t] += gz;
                gZ1[mb][t] = (short)f2bf(gz);
            }
        }
        // dW2[k][m] = sum_s H1[s][k] gZ2[s][m]: operands transposed through LDS
        //   A (16x16 block kb): lane: H1^T[16kb + i][sample 4q+t]
        //   B (16x16 block mb): lane: gZ2^T[16mb + i][sample 4q+t]
        short4_t AH1[4], BG2[4], BG1[4];
#pragma unroll
        for (int b = 0; b < 4; b++) {
#pragma unroll
            for (int t = 0; t < 4; t++) {
                tr[wid][0][4 * q + t][i] = bf2f((unsigned short)T.H1[b][t]);
                tr[wid][1][4 * q + t][i] = bf2f((unsigned short)gZ2[b][t]);
            }
            wave_lds_sync();
#pragma unroll
            for (int t = 0; t < 4; t++) {
                AH1[b][t] = (short)f2bf(tr[wid][0][i][4 * q + t]);
                BG2[b][t] = (short)f2bf(tr[wid][1][i][4 * q + t]);
            }
            wave_lds_sync();
#pragma unroll
            for (int t = 0; t < 4; t++) tr[wid][1][4 * q + t][i] = bf2f((unsigned short)gZ1[b][t]);
            wave_lds_sync();
#pragma unroll
            for (int t = 0; t < 4; t++) BG1[b][t] = (short)f2bf(tr[wid][1][i][4 * q + t]);
            wave_lds_sync();
        }
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
#pragma unroll
            for (int mb = 0; mb < 4; mb++) gW2[kb * 4 + mb] = mfma(AH1[kb], BG2[mb], gW2[kb * 4 + mb]);
        // dW1[f][k] = sum_s X[s][f] gZ1[s][k]; A lane: X[sample 4q+t][feature i]
#pragma unroll
        for (int mb = 0; mb < 4; mb++) gW1[mb] = mfma(AX, BG1[mb], gW1[mb]);
    }

    // ---- reduce the wave accumulators into one partial vector per work-group
    // Layouts of the accumulators (lane l = 16q + i):
    //   gW2[kb*4+mb][t]: dW2[16kb + 4q + t][16mb + i]
    //   gW1[mb][t]:      dW1[4q + t][16mb + i]
    //   gW3[mb][t]:      dW3[16mb + 4q + t] (to be summed over i)
    //   gb2/gb1[mb][t]:  db[16mb + 4q + t]    (to be summed over i)
    float *out = partial + (long)blockIdx.x * NPARAM_PAD;
    // sum over the 16 lanes that share q (the samples of a tile)
    auto sum16 = [](float v) {
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 16);
        return v;
    };
    // direct form: every lane writes its values into a per-wave LDS image of
    // the parameter vector, then the 4 waves are summed
    __shared__ float img[NPARAM_PAD];
    for (int p = threadIdx.x; p < NPARAM_PAD; p += GB) img[p] = 0.f;
    __syncthreads();
    for (int w = 0; w < GB / 64; w++) {
        if (wid == w) {
#pragma unroll
            for (int kb = 0; kb < 4; kb++)
#pragma unroll
                for (int mb = 0; mb < 4; mb++)
#pragma unroll
                    for (int t = 0; t < 4; t++) img[P_W2 + (16 * kb + 4 * q + t) * NH + 16 * mb + i] += gW2[kb * 4 + mb][t];
#pragma unroll
            for (int mb = 0; mb < 4; mb++)
#pragma unroll
                for (int t = 0; t < 4; t++) img[P_W1 + (4 * q + t) * NH + 16 * mb + i] += gW1[mb][t];
#pragma unroll
            for (int mb = 0; mb < 4; mb++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const float w3 = sum16(gW3[mb][t]);
                    const float b2 = sum16(gb2[mb][t]);
                    const float b1 = sum16(gb1[mb][t]);
                    if (i == 0) {
                        img[P_W3 + 16 * mb + 4 * q + t] += w3;
                        img[P_B2 + 16 * mb + 4 * q + t] += b2;
                        img[P_B1 + 16 * mb + 4 * q + t] += b1;
                    }
                }
            float b3 = gb3, ls = loss;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                b3 += __shfl_xor(b3, off, 64);
                ls += __shfl_xor(ls, off, 64);
            }
            if (l == 0) {
                img[P_B3] += b3;
                img[NPARAM] += ls;  // loss slot (inside the padding)
            }
        }
        __syncthreads();
    }
    for (int p = threadIdx.x; p < NPARAM_PAD; p += GB) out[p] = img[p];
}

// sum the work-group partials (deterministic order) and take one Adam step
// The gradient is the sum of the nblocks per-work-group partials: each
// work-group takes 64 parameters, its four wavefronts a quarter of the
// partials each (four running sums per lane, so the loads overlap), in a
// fixed order; then the Adam update.
__global__ void __launch_bounds__(GB) k_adam(int nblocks, const float *__restrict__ partial, float *__restrict__ P,
                                             float *__restrict__ m, float *__restrict__ v, float *__restrict__ grad_out,
                                             float lr, float b1, float b2, float eps, int step)
{
    __shared__ float sm[GB / 64][64];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    const int b0 = (int)((long)nblocks * q / (GB / 64)), b1e = (int)((long)nblocks * (q + 1) / (GB / 64));
    float a[4] = {0.f, 0.f, 0.f, 0.f};
    int b = b0;
    for (; b + 4 <= b1e; b += 4)
#pragma unroll
        for (int t = 0; t < 4; t++) a[t] += partial[(long)(b + t) * NPARAM_PAD + p];
    for (; b < b1e; b++) a[0] += partial[(long)b * NPARAM_PAD + p];
    sm[q][lane] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    if (q != 0) return;
    const float g = ((sm[0][lane] + sm[1][lane]) + sm[2][lane]) + sm[3][lane];
    if (grad_out) grad_out[p] = g;
    if (p >= NPARAM || lr <= 0.f) return;
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.f - powf(b1, (float)step));
    const float vh = vv / (1.f - powf(b2, (float)step));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------------------
// features of every row, packed as BF16 [rows][16]
__global__ void __launch_bounds__(GB) k_features(int rows, int n, const double *__restrict__ r0,
                                                 const double *__restrict__ Ar0, const double *__restrict__ A2r0,
                                                 const double *__restrict__ H, const double *__restrict__ AH, int s1,
                                                 int s2, const uint8_t *__restrict__ dirichlet,
                                                 const double *__restrict__ scal, unsigned short *__restrict__ X)
{
    const int r = blockIdx.x * GB + threadIdx.x;
    if (r >= rows) return;
    // the newest increment (slot s1), the one before it (s2) and A times the newest
    const double dx1 = H[(long)r * HIST_SLOTS + s1], dx2 = H[(long)r * HIST_SLOTS + s2];
    const double Adx1 = AH[(long)r * HIST_SLOTS + s1];
    const double isr = scal[0] > 0 ? 1.0 / scal[0] : 0.0;
    const double isx = scal[1] > 0 ? 1.0 / scal[1] : 0.0;
    float f[NF] = {};
    f[0] = (float)(r0[r] * isr);
    f[1] = (float)(Ar0[r] * isr);
    f[2] = (float)(A2r0[r] * isr);
    f[3] = (float)(dx1 * isx);
    f[4] = (float)(dx2 * isx);
    f[5] = (float)(Adx1 * isx);
    f[6] = dirichlet[r] ? 1.f : 0.f;
    f[7] = n > 1 ? (float)(r % n) / (float)(n - 1) : 0.f;
    f[8] = (scal[0] > 0 && scal[1] > 0) ? (float)(log10(scal[0] / scal[1]) * 0.25) : 0.f;
    unsigned short *o = X + (long)r * NF;
#pragma unroll
    for (int k = 0; k < NF; k++) o[k] = f2bf(f[k]);
}

// direction predicted by the network: d = s_x * y
__global__ void __launch_bounds__(GB) k_direction(int rows, const float *__restrict__ y, const double *__restrict__ scal,
                                                  double *__restrict__ d)
{
    const int r = blockIdx.x * GB + threadIdx.x;
    if (r < rows) d[r] = scal[1] * (double)y[r];
}

constexpr int HMAX = HIST_SLOTS;    // previous increments kept
constexpr int MAXDIR = HMAX + 1;     // + the network's direction

// Work-group sum (every lane calls it); the result is returned to lane 0.
__device__ __forceinline__ double group_sum(double v)
{
    __shared__ double sm[GB / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
#pragma unroll
        for (int w = 0; w < GB / 64; w++) t += sm[w];
    return t;
}

// squared norm of v, one partial per work-group
__global__ void __launch_bounds__(GB) k_norm_part(int rows, const double *__restrict__ v, double *__restrict__ part)
{
    const int r = blockIdx.x * GB + threadIdx.x;
    const double x = r < rows ? v[r] : 0.0;
    const double t = group_sum(x * x);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// The history: H[r*HIST_SLOTS + s] holds increment s of the ring (FP64) and
// AH the same slot of A times it; a row's slots sit together, so the kernels
// that use every slot (the product, the Gram matrix, the combination) read a
// row in one piece.
//
// x = x_prev + a_net d + sum_s a_s H[., s]; G4 groups of 4 slots in use.
struct HistCoef {
    double a[HIST_SLOTS];
    double a_net;
};
template <int G4>
__global__ void __launch_bounds__(GB) k_combine_hist(int rows, const double *__restrict__ xp,
                                                     const double *__restrict__ H, const double *__restrict__ dnet,
                                                     HistCoef K, double *__restrict__ x)
{
    const int r = blockIdx.x * GB + threadIdx.x;
    if (r >= rows) return;
    double v = xp[r];
    if (dnet) v += K.a_net * dnet[r];
#pragma unroll
    for (int g = 0; g < 2 * G4; g++) {
        const Pack<double, 2> h = ld<double, 2>(H + (long)r * HIST_SLOTS + 2 * g);
        v += K.a[2 * g] * h.v[0];
        v += K.a[2 * g + 1] * h.v[1];
    }
    x[r] = v;
}

// the increment of this solve becomes the newest history entry (and, while
// the network trains, its normalised target); the partials of its squared
// norm (as stored) give the next solve's feature scale
__global__ void __launch_bounds__(GB) k_after_solve(int rows, const double *__restrict__ x, const double *__restrict__ xp,
                                                    double *__restrict__ H, int slot, const double *__restrict__ scal,
                                                    float *__restrict__ target, double *__restrict__ part)
{
    const int r = blockIdx.x * GB + threadIdx.x;
    double e = 0.0;
    if (r < rows) {
        const double d = x[r] - xp[r];
        if (target) target[r] = scal[1] > 0 ? (float)(d / scal[1]) : 0.f;
        H[(long)r * HIST_SLOTS + slot] = d;
        e = d * d;
    }
    const double t = group_sum(e);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Gram matrix of u = [r0, w, AH slot 0 .. HIST_SLOTS-1] (w: A times the
// network's direction, zero when absent): all NPROD products u_p . u_q,
// q <= p, packed at p(p+1)/2 + q, in one pass.  The four wavefronts of a
// work-group read the same 64 rows and each forms a quarter of the products
// (triangle rows 0-8, 9-12, 13-15, 16-17: 45, 46, 45 and 35 products), so
// the vectors come from memory once.
// One partial per work-group and product; k_gram_final sums them in a fixed
// order, so the result does not depend on timing.
constexpr int NU = 2 + HIST_SLOTS, NPROD = NU * (NU + 1) / 2;
constexpr int GACC = 46;
template <int P0, int P1>
__device__ __forceinline__ void gram_rows(const double (&u)[NU], double (&acc)[GACC])
{
#pragma unroll
    for (int p = P0; p < P1; p++)
#pragma unroll
        for (int q = 0; q <= p; q++) acc[p * (p + 1) / 2 + q - P0 * (P0 + 1) / 2] += u[p] * u[q];
}
template <int P0, int P1>
__device__ __forceinline__ void gram_store(const double (&acc)[GACC], double *__restrict__ partial)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int p = P0; p < P1; p++)
#pragma unroll
        for (int q = 0; q <= p; q++) {
            double v = acc[p * (p + 1) / 2 + q - P0 * (P0 + 1) / 2];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) partial[(long)(p * (p + 1) / 2 + q) * gridDim.x + blockIdx.x] = v;
        }
}
__global__ void __launch_bounds__(GB) k_gram_hist(int rows, const double *__restrict__ r0, const double *__restrict__ w,
                                                  const double *__restrict__ AH, double *__restrict__ partial)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double acc[GACC] = {};
    for (int r = blockIdx.x * 64 + lane; r < rows; r += gridDim.x * 64) {
        double u[NU];
        u[0] = r0[r];
        u[1] = w ? w[r] : 0.0;
#pragma unroll
        for (int g = 0; g < HIST_SLOTS / 2; g++) {
            const Pack<double, 2> h = ld<double, 2>(AH + (long)r * HIST_SLOTS + 2 * g);
            u[2 + 2 * g] = h.v[0];
            u[3 + 2 * g] = h.v[1];
        }
        switch (wid) {
        case 0: gram_rows<0, 9>(u, acc); break;
        case 1: gram_rows<9, 13>(u, acc); break;
        case 2: gram_rows<13, 16>(u, acc); break;
        default: gram_rows<16, 18>(u, acc); break;
        }
    }
    switch (wid) {
    case 0: gram_store<0, 9>(acc, partial); break;
    case 1: gram_store<9, 13>(acc, partial); break;
    case 2: gram_store<13, 16>(acc, partial); break;
    default: gram_store<16, 18>(acc, partial); break;
    }
}

// One work-group per product: the nb partials summed in a fixed order.
__global__ void __launch_bounds__(GB) k_gram_final(int nb, const double *__restrict__ partial, double *__restrict__ out)
{
    const int q = blockIdx.x;
    double v = 0;
    for (int b = threadIdx.x; b < nb; b += GB) v += partial[(long)q * nb + b];
    const double t = group_sum(v);
    if (threadIdx.x == 0) out[q] = t;
}

}  // namespace mhfem

// ===========================================================================
// host side
namespace mhfem {

struct Guess {
    mhfem_guess_config cfg{};
    int rows = 0;
    DBuf<unsigned short> X;
    DBuf<float> y, target;
    DBuf<double> xp, r0, Ar0, A2r0, rtry, scal;
    DBuf<double> H, AH;     // previous increments (ring of `depth` slots) and A times them
    DBuf<double> norm_part, norms;  // squared norms of r0 and of the newest increment
    int head = 0;           // ring slot of the newest increment
    int slot(int j) const { return (head + j) % depth; }  // slot of the j-th newest
    DBuf<uint8_t> dirichlet;
    DBuf<float> P, m, v, partial, grad;
    DBuf<PackedWeights> W;
    DBuf<double> gram_part, gram;
    double *h_gram = nullptr;  // pinned, NPROD values
    int gram_blocks = 0;
    int nblocks = 0;
    int adam_step = 0;
    int solves = 0;        // solves seen
    int trained = 0;       // solves used for training
    int history = 0;       // increments available (0 .. depth)
    int depth = 2;         // increments used (history_depth)
    bool have_weights = false;
    int net_fail = 0;      // consecutive solves in which the network did not help
    bool net_off = false;  // switched off by the reject_after rule
    bool features = false; // X holds the features of the current solve
    mhfem_guess_stats last{};
};

namespace {

inline unsigned gblocks(long n) { return (unsigned)((n + GB - 1) / GB); }

void init_params(Guess &G)
{
    // He-normal initialisation of the hidden layers, zero output layer
    std::vector<float> P(NPARAM_PAD, 0.f);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ (uint64_t)G.cfg.seed;
    auto unif = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto normal = [&]() { return std::sqrt(-2.0 * std::log(unif())) * std::cos(6.283185307179586 * unif()); };
    for (int k = 0; k < NF * NH; k++) P[P_W1 + k] = (float)(normal() * std::sqrt(2.0 / NF));
    for (int k = 0; k < NH * NH; k++) P[P_W2 + k] = (float)(normal() * std::sqrt(2.0 / NH));
    MHFEM_HIP(hipMemcpy(G.P.p, P.data(), NPARAM_PAD * sizeof(float), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemset(G.m.p, 0, NPARAM_PAD * sizeof(float)));
    MHFEM_HIP(hipMemset(G.v.p, 0, NPARAM_PAD * sizeof(float)));
    G.adam_step = 0;
}

void pack(Ctx &c, Guess &G)
{
    hipLaunchKernelGGL(k_pack_weights, 1, 64, 0, c.stream, G.P.p, G.W.p);
    MHFEM_HIP(hipGetLastError());
}


// one full-batch Adam step on (X, target); returns the mean squared loss
double train_epoch(Ctx &c, Guess &G, long rows, float lr)
{
    const long ntiles = (rows + 15) / 16;
    const long stride = std::max(1, G.cfg.sample_stride);
    const long offset = G.adam_step % stride;
    const long nsub = (ntiles - offset + stride - 1) / stride;
    hipLaunchKernelGGL(k_mlp_grad, G.nblocks, GB, 0, c.stream, rows, G.X.p, G.target.p, G.W.p,
                       1.0f / (float)(nsub * 16), G.partial.p, stride, offset);
    G.adam_step++;
    hipLaunchKernelGGL(k_adam, NPARAM_PAD / 64, GB, 0, c.stream, G.nblocks, G.partial.p, G.P.p, G.m.p,
                       G.v.p, G.grad.p, lr, 0.9f, 0.999f, 1e-8f, G.adam_step);
    MHFEM_HIP(hipGetLastError());
    pack(c, G);
    return -1.0;
}

}  // namespace

Guess *guess_create(Ctx &c)
{
    Guess *G = new Guess();
    const int n = c.rows;
    G->rows = n;
    G->X.alloc((size_t)n * NF);
    G->y.alloc(n);
    G->target.alloc(n);
    for (DBuf<double> *b : {&G->xp, &G->r0, &G->Ar0, &G->A2r0, &G->rtry}) {
        b->alloc(n);
        MHFEM_HIP(hipMemset(b->p, 0, (size_t)n * sizeof(double)));
    }
    for (DBuf<double> *b : {&G->H, &G->AH}) {
        b->alloc((size_t)n * HIST_SLOTS);
        MHFEM_HIP(hipMemset(b->p, 0, (size_t)n * HIST_SLOTS * sizeof(double)));
    }
    G->norm_part.alloc(2 * (size_t)gblocks(n));
    MHFEM_HIP(hipMemset(G->norm_part.p, 0, 2 * (size_t)gblocks(n) * sizeof(double)));
    G->norms.alloc(2);
    G->scal.alloc(4);
    std::vector<uint8_t> dflag(n);
    for (int r = 0; r < n; r++) dflag[r] = (c.h_rowptr[r + 1] - c.h_rowptr[r]) == 1;
    G->dirichlet.alloc(n);
    MHFEM_HIP(hipMemcpy(G->dirichlet.p, dflag.data(), n, hipMemcpyHostToDevice));
    G->P.alloc(NPARAM_PAD);
    G->m.alloc(NPARAM_PAD);
    G->v.alloc(NPARAM_PAD);
    G->grad.alloc(NPARAM_PAD);
    G->W.alloc(1);
    const long ntiles = ((long)n + 15) / 16;
    G->nblocks = (int)std::max<long>(1, std::min<long>(1024, (ntiles + 3) / 4));
    G->partial.alloc((size_t)G->nblocks * NPARAM_PAD);
    G->gram_blocks = (int)std::max<long>(1, std::min<long>(1024, ((long)n + GB - 1) / GB));
    G->gram_part.alloc((size_t)G->gram_blocks * NPROD);
    G->gram.alloc(NPROD);
    MHFEM_HIP(hipHostMalloc((void **)&G->h_gram, NPROD * sizeof(double), hipHostMallocDefault));
    G->cfg.enabled = 0;
    G->cfg.train_steps = 100;
    G->cfg.epochs = 20;
    G->cfg.learning_rate = 1e-3f;
    G->cfg.seed = 1;
    init_params(*G);
    pack(c, *G);
    return G;
}

void guess_destroy(Guess *G)
{
    if (G && G->h_gram) (void)hipHostFree(G->h_gram);
    delete G;
}

void guess_configure(Ctx &c, Guess &G, const mhfem_guess_config &cfg)
{
    const bool reseed = cfg.seed != G.cfg.seed;
    G.cfg = cfg;
    if (G.cfg.epochs < 0) G.cfg.epochs = 0;
    G.net_fail = 0;
    G.net_off = false;
    const int depth = cfg.history_depth > 0 ? std::min(cfg.history_depth, HMAX) : 2;
    if (depth != G.depth) {
        // a new depth starts an empty history in a ring of `depth` slots
        G.depth = depth;
        G.history = 0;
        G.head = 0;
        for (DBuf<double> *b : {&G.H, &G.AH})
            MHFEM_HIP(hipMemset(b->p, 0, (size_t)G.rows * HIST_SLOTS * sizeof(double)));
        MHFEM_HIP(hipMemset(G.norm_part.p, 0, G.norm_part.n * sizeof(double)));
    }
    if (reseed && !G.have_weights) {
        init_params(G);
        pack(c, G);
    }
}

namespace {

// All products of u = [r0, w, AH slots] (k_gram_hist), read back with one
// synchronisation into G.h_gram; gram_at(p, q) unpacks them.
void gram_hist(Ctx &c, Guess &G, const double *w)
{
    hipLaunchKernelGGL(k_gram_hist, G.gram_blocks, GB, 0, c.stream, G.rows, G.r0.p, w, G.AH.p, G.gram_part.p);
    hipLaunchKernelGGL(k_gram_final, NPROD, GB, 0, c.stream, G.gram_blocks, G.gram_part.p, G.gram.p);
    MHFEM_HIP(hipGetLastError());
    MHFEM_HIP(hipMemcpyAsync(G.h_gram, G.gram.p, NPROD * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    MHFEM_HIP(hipStreamSynchronize(c.stream));
}

double gram_at(const Guess &G, int p, int q)
{
    if (q > p) std::swap(p, q);
    return G.h_gram[p * (p + 1) / 2 + q];
}

// min || r0 - sum_{p in [lo, k)} a_p A D_p || by the normal equations
// Gm a = gv.  The directions are scaled to unit norm and the Gram matrix is
// factorised by Cholesky with diagonal pivoting, stopping at pivots below
// 1e-13: directions (nearly) in the span of the chosen ones get a = 0, so
// that nearly parallel previous increments do not blow the coefficients up.
// Returns the predicted squared residual rr - 2 a.gv + a.Gm.a.
double least_residual(const double (&Gm)[MAXDIR][MAXDIR], const double (&gv)[MAXDIR], double rr, int lo, int k,
                      double (&a)[MAXDIR])
{
    for (int p = 0; p < MAXDIR; p++) a[p] = 0.0;
    const int m = k - lo;
    if (m <= 0) return rr;
    double sc[MAXDIR], S[MAXDIR][MAXDIR], L[MAXDIR][MAXDIR] = {}, dg[MAXDIR];
    for (int p = 0; p < m; p++) sc[p] = Gm[lo + p][lo + p] > 0 ? 1.0 / std::sqrt(Gm[lo + p][lo + p]) : 0.0;
    for (int p = 0; p < m; p++)
        for (int q = 0; q < m; q++) S[p][q] = Gm[lo + p][lo + q] * sc[p] * sc[q];
    for (int p = 0; p < m; p++) dg[p] = S[p][p];
    int piv[MAXDIR], used = 0;
    bool taken[MAXDIR] = {};
    for (int step = 0; step < m; step++) {
        int j = -1;
        for (int p = 0; p < m; p++)
            if (!taken[p] && (j < 0 || dg[p] > dg[j])) j = p;
        if (j < 0 || !(dg[j] > 1e-13)) break;
        taken[j] = true;
        piv[used] = j;
        L[j][used] = std::sqrt(dg[j]);
        for (int p = 0; p < m; p++) {
            if (taken[p]) continue;
            double v = S[p][j];
            for (int t = 0; t < used; t++) v -= L[p][t] * L[j][t];
            L[p][used] = v / L[j][used];
            dg[p] -= L[p][used] * L[p][used];
        }
        used++;
    }
    if (used == 0) return rr;
    // solve (L_pp L_pp^T) z = gs on the chosen directions
    double y[MAXDIR], z[MAXDIR];
    for (int t = 0; t < used; t++) {
        double v = gv[lo + piv[t]] * sc[piv[t]];
        for (int u = 0; u < t; u++) v -= L[piv[t]][u] * y[u];
        y[t] = v / L[piv[t]][t];
    }
    for (int t = used - 1; t >= 0; t--) {
        double v = y[t];
        for (int u = t + 1; u < used; u++) v -= L[piv[u]][t] * z[u];
        z[t] = v / L[piv[t]][t];
    }
    for (int t = 0; t < used; t++) a[lo + piv[t]] = z[t] * sc[piv[t]];
    double pred = rr;
    for (int p = lo; p < k; p++) {
        pred -= 2 * a[p] * gv[p];
        for (int q = lo; q < k; q++) pred += a[p] * Gm[p][q] * a[q];
    }
    return pred;
}

}  // namespace

// Before a solve: features of the warm start and, once the network has been
// trained, the predicted starting vector (kept only if it lowers the residual).
void guess_before_solve(Ctx &c, Guess &G, const double *b, double *x)
{
    const int n = G.rows;
    const unsigned g = gblocks(n);
    G.last = mhfem_guess_stats{};
    G.last.network_active = !G.net_off;
    G.features = false;
    MHFEM_HIP(hipMemcpyAsync(G.xp.p, x, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    residual64(c, x, b, G.r0.p);
    // |r0|^2 and |newest increment|^2 (its partials come from the last
    // k_after_solve), read back with one synchronisation
    hipLaunchKernelGGL(k_norm_part, g, GB, 0, c.stream, n, G.r0.p, G.norm_part.p);
    hipLaunchKernelGGL(k_gram_final, 2, GB, 0, c.stream, (int)g, G.norm_part.p, G.norms.p);
    MHFEM_HIP(hipGetLastError());
    MHFEM_HIP(hipMemcpyAsync(G.h_gram, G.norms.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    MHFEM_HIP(hipStreamSynchronize(c.stream));
    const double rr = G.h_gram[0];
    const double sx2 = G.h_gram[1] / n;
    G.last.r0_norm = std::sqrt(rr);
    G.last.r_guess_norm = G.last.r0_norm;
    if (!(G.cfg.enabled) || G.history < 1 || rr == 0.0) return;
    // The network is evaluated while it trains (its features pair with this
    // solve's increment) and afterwards until the reject_after rule switches
    // it off.
    const bool trained = G.have_weights || G.trained > 0;
    const bool training = G.cfg.epochs > 0 && G.trained < G.cfg.train_steps;
    const bool want_net = G.cfg.combine != MHFEM_GUESS_HISTORY && !G.net_off && sx2 > 0.0;
    const bool use_net = want_net && trained;
    const bool use_hist = G.cfg.combine != MHFEM_GUESS_NETWORK && sx2 > 0.0;
    const int h = std::min(G.history, G.depth);
    if (sx2 > 0.0) {
        double sc[4] = {std::sqrt(rr / n), std::sqrt(sx2), 0, 0};
        MHFEM_HIP(hipMemcpyAsync(G.scal.p, sc, sizeof(sc), hipMemcpyHostToDevice, c.stream));
        // A times every slot of the ring, one pass over the matrix
        spmm_hist(c, G.H.p, G.AH.p, G.depth);
    }
    if (want_net && (use_net || training)) {
        spmv64(c, G.r0.p, G.Ar0.p);
        spmv64(c, G.Ar0.p, G.A2r0.p);
        const int s2 = G.depth > 1 ? G.slot(1) : G.slot(0);
        hipLaunchKernelGGL(k_features, g, GB, 0, c.stream, n, c.n, G.r0.p, G.Ar0.p, G.A2r0.p, G.H.p, G.AH.p,
                           G.slot(0), s2, G.dirichlet.p, G.scal.p, G.X.p);
        MHFEM_HIP(hipGetLastError());
        G.features = true;
    }
    // Directions: the network's predicted increment and (unless the network
    // runs alone) the previous increments.  The starting vector is the
    // combination x_prev + sum a_k D_k of least residual norm, so it is never
    // worse than the warm start; it is kept only if it lowers the residual.
    // Direction k is vector u[k] of the Gram matrix: u1 = A d for the
    // network's direction d, u(2 + s) = A times ring slot s.
    int uidx[MAXDIR];
    int k = 0;
    if (use_net) {
        const long ntiles = ((long)n + 15) / 16;
        hipLaunchKernelGGL(k_mlp_forward, (unsigned)std::max<long>(1, std::min<long>(4096, (ntiles + 3) / 4)), GB, 0,
                           c.stream, (long)n, G.X.p, G.W.p, G.y.p);
        hipLaunchKernelGGL(k_direction, g, GB, 0, c.stream, n, G.y.p, G.scal.p, G.rtry.p);
        MHFEM_HIP(hipGetLastError());
        spmv64(c, G.rtry.p, G.Ar0.p);  // Ar0 is free once the features are packed
        uidx[k++] = 1;
    }
    if (use_hist)
        for (int j = 0; j < h; j++) uidx[k++] = 2 + G.slot(j);
    if (k == 0) return;
    gram_hist(c, G, use_net ? G.Ar0.p : nullptr);
    double Gm[MAXDIR][MAXDIR] = {}, gv[MAXDIR] = {}, a[MAXDIR] = {};
    for (int p = 0; p < k; p++) {
        gv[p] = gram_at(G, uidx[p], 0);
        for (int q = 0; q <= p; q++) Gm[p][q] = Gm[q][p] = gram_at(G, uidx[p], uidx[q]);
    }
    const double pred = least_residual(Gm, gv, rr, 0, k, a);
    G.last.r_guess_norm = std::sqrt(std::max(pred, 0.0));
    G.last.coef_network = use_net ? a[0] : 0.0;
    if (use_net && use_hist && k > 1) {
        // what the network adds to the previous increments alone
        double ah[MAXDIR];
        const double pred_hist = least_residual(Gm, gv, rr, 1, k, ah);
        G.last.network_gain = pred_hist > 0 ? (pred_hist - pred) / pred_hist : 0.0;
        if (!training && G.cfg.reject_after > 0) {
            G.net_fail = G.last.network_gain < 0.01 ? G.net_fail + 1 : 0;
            if (G.net_fail >= G.cfg.reject_after) G.net_off = true;
        }
    }
    if (!(pred < rr)) return;
    HistCoef K{};
    for (int p = 0; p < k; p++) {
        if (uidx[p] == 1) K.a_net = a[p];
        else K.a[uidx[p] - 2] = a[p];
    }
    const double *dnet = use_net ? G.rtry.p : nullptr;
    switch ((G.depth + 3) / 4) {
    case 1: hipLaunchKernelGGL(k_combine_hist<1>, g, GB, 0, c.stream, n, G.xp.p, G.H.p, dnet, K, x); break;
    case 2: hipLaunchKernelGGL(k_combine_hist<2>, g, GB, 0, c.stream, n, G.xp.p, G.H.p, dnet, K, x); break;
    case 3: hipLaunchKernelGGL(k_combine_hist<3>, g, GB, 0, c.stream, n, G.xp.p, G.H.p, dnet, K, x); break;
    default: hipLaunchKernelGGL(k_combine_hist<4>, g, GB, 0, c.stream, n, G.xp.p, G.H.p, dnet, K, x); break;
    }
    MHFEM_HIP(hipGetLastError());
    G.last.used = 1;
}

// After a solve: the increment becomes the newest history entry and, during
// the training phase, a training target.
void guess_after_solve(Ctx &c, Guess &G, const double *x)
{
    const int n = G.rows;
    const unsigned g = gblocks(n);
    const bool train = G.cfg.enabled && G.features && G.trained < G.cfg.train_steps && G.cfg.epochs > 0;
    // the newest increment takes the slot of the oldest (this solve's
    // features are already packed in X)
    G.head = (G.head + G.depth - 1) % G.depth;
    hipLaunchKernelGGL(k_after_solve, g, GB, 0, c.stream, n, x, G.xp.p, G.H.p, G.slot(0), G.scal.p,
                       train ? G.target.p : (float *)nullptr, G.norm_part.p + g);
    MHFEM_HIP(hipGetLastError());
    if (train) {
        for (int e = 0; e < G.cfg.epochs; e++) train_epoch(c, G, n, G.cfg.learning_rate);
        G.trained++;
        // loss of the last epoch
        float loss = 0.f;
        MHFEM_HIP(hipMemcpyAsync(&loss, G.grad.p + NPARAM, sizeof(float), hipMemcpyDeviceToHost, c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
        G.last.loss = loss / n;
    }
    G.last.trained_steps = G.trained;
    G.history = std::min(G.history + 1, G.depth);
    G.solves++;
}

void guess_forward(Ctx &c, Guess &G, const uint16_t *X, long rows, float *y)
{
    DBuf<unsigned short> dX;
    DBuf<float> dy;
    dX.alloc((size_t)rows * NF);
    dy.alloc(rows);
    MHFEM_HIP(hipMemcpy(dX.p, X, (size_t)rows * NF * 2, hipMemcpyHostToDevice));
    const long ntiles = (rows + 15) / 16;
    hipLaunchKernelGGL(k_mlp_forward, (unsigned)std::max<long>(1, std::min<long>(4096, (ntiles + 3) / 4)), GB, 0,
                       c.stream, rows, dX.p, G.W.p, dy.p);
    MHFEM_HIP(hipGetLastError());
    MHFEM_HIP(hipMemcpyAsync(y, dy.p, rows * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    MHFEM_HIP(hipStreamSynchronize(c.stream));
}

void guess_gradient(Ctx &c, Guess &G, const uint16_t *X, const float *t, long rows, float *grad)
{
    DBuf<unsigned short> dX;
    DBuf<float> dt;
    dX.alloc((size_t)rows * NF);
    dt.alloc(rows);
    MHFEM_HIP(hipMemcpy(dX.p, X, (size_t)rows * NF * 2, hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(dt.p, t, rows * sizeof(float), hipMemcpyHostToDevice));
    const long ntiles = (rows + 15) / 16;
    const int nb = (int)std::max<long>(1, std::min<long>(G.nblocks, (ntiles + 3) / 4));
    hipLaunchKernelGGL(k_mlp_grad, nb, GB, 0, c.stream, rows, dX.p, dt.p, G.W.p, 1.0f / (float)rows, G.partial.p,
                       1L, 0L);
    // reduce only (learning rate 0)
    hipLaunchKernelGGL(k_adam, NPARAM_PAD / 64, GB, 0, c.stream, nb, G.partial.p, G.P.p, G.m.p, G.v.p,
                       G.grad.p, 0.0f, 0.9f, 0.999f, 1e-8f, 1);
    MHFEM_HIP(hipGetLastError());
    MHFEM_HIP(hipMemcpyAsync(grad, G.grad.p, (NPARAM + 1) * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    MHFEM_HIP(hipStreamSynchronize(c.stream));
}

void guess_get_weights(Ctx &c, Guess &G, float *out)
{
    MHFEM_HIP(hipMemcpyAsync(out, G.P.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    MHFEM_HIP(hipStreamSynchronize(c.stream));
}

void guess_set_weights(Ctx &c, Guess &G, const float *in)
{
    MHFEM_HIP(hipMemcpyAsync(G.P.p, in, NPARAM * sizeof(float), hipMemcpyHostToDevice, c.stream));
    pack(c, G);
    MHFEM_HIP(hipStreamSynchronize(c.stream));
    G.have_weights = true;
}

int guess_num_params() { return NPARAM; }

mhfem_guess_stats guess_last_stats(const Guess &G) { return G.last; }

}  // namespace mhfem
