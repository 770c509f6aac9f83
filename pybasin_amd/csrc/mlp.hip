// mlp.hip -- kernels of the learned surrogate: IC (theta, omega, T) -> attractor.
//
// Network: 3 -> 256 -> 256 -> 256 -> 2 with SiLU, weights and activations in
// bfloat16, products accumulated in float32 on the matrix cores
// (v_mfma_f32_16x16x16bf16_1k).  The host composes these kernels into the
// forward pass, the backward pass (cross-entropy loss) and the Adam update;
// the surrogate has no counterpart in the reference -- it is trained on
// labels the Dopri5 integrator produced.
//
// Layouts: activations row-major (rows = samples); a layer computes
// Z = A . W + b with W stored (in, out) row-major.  The input layer is padded
// from 3 to 16 columns (zeros) to fit the 16-deep product.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace bh {
namespace mlp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (NaN kept quiet)
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }

// C[m, n] (+)= sum_k A(m, k) B(k, n); A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]
// (bf16 bits), C float32 with row stride ldc.  Block = 4 waves, block tile
// 64 x 64, each wave a 32 x 32 tile of four 16x16 products; K in steps of 16.
struct GemmArgs {
    const unsigned short *A;
    const unsigned short *B;
    float *C;
    int64_t M, N, K;
    int64_t sam, sak, sbk, sbn, ldc;
    int accumulate;
};

__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t m0 = (int64_t)blockIdx.y * 64 + (wave >> 1) * 32;
    const int64_t n0 = (int64_t)blockIdx.x * 64 + (wave & 1) * 32;
    const int li = lane & 15, lk = (lane >> 4) * 4;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = (floatx4){0.f, 0.f, 0.f, 0.f};
    for (int64_t k0 = 0; k0 < g.K; k0 += 16) {
        bf16x4 af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int64_t m = m0 + 16 * i + li;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int64_t k = k0 + lk + e;
                af[i][e] = (m < g.M && k < g.K) ? (short)g.A[m * g.sam + k * g.sak] : (short)0;
            }
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int64_t n = n0 + 16 * j + li;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int64_t k = k0 + lk + e;
                bfr[j][e] = (n < g.N && k < g.K) ? (short)g.B[k * g.sbk + n * g.sbn] : (short)0;
            }
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int64_t n = n0 + 16 * j + li;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t m = m0 + 16 * i + (lane >> 4) * 4 + r;
                if (m < g.M && n < g.N) {
                    float *c = g.C + m * g.ldc + n;
                    *c = g.accumulate ? *c + acc[i][j][r] : acc[i][j][r];
                }
            }
        }
}

// Z[m, n] += b[n]; optionally keep Z (pre-activation) in zsave; write
// act(Z) (SiLU when silu != 0) as bf16 to out[m * ldo + n].
__global__ void bias_act_kernel(const float *Z, int64_t M, int64_t N, const float *b, int silu,
                                float *zsave, unsigned short *out, int64_t ldo, float *out_f32) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * N) return;
    const int64_t m = idx / N, n = idx % N;
    const float z = Z[idx] + b[n];
    if (zsave) zsave[idx] = z;
    const float a = silu ? z / (1.0f + __expf(-z)) : z;
    if (out) out[m * ldo + n] = f2bf(a);
    if (out_f32) out_f32[idx] = a;
}

// dZ = dA * silu'(z), silu'(z) = s (1 + z (1 - s)), s = sigmoid(z); f32 and bf16 copies
__global__ void silu_bwd_kernel(const float *dA, const float *Z, int64_t count, float *dZ,
                                unsigned short *dZb) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const float z = Z[idx];
    const float s = 1.0f / (1.0f + __expf(-z));
    const float g = dA[idx] * (s * (1.0f + z * (1.0f - s)));
    dZ[idx] = g;
    dZb[idx] = f2bf(g);
}

// two-class softmax cross-entropy on logits (M x 2, f32, bias already added):
// dlogits = (softmax - onehot) / M (f32 and bf16), per-row loss
__global__ void softmax_ce_kernel(const float *logits, const int32_t *labels, int64_t M,
                                  float *dlog, unsigned short *dlogb, float *loss) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const float l0 = logits[2 * m], l1 = logits[2 * m + 1];
    const float mx = fmaxf(l0, l1);
    const float e0 = __expf(l0 - mx), e1 = __expf(l1 - mx);
    const float s = e0 + e1;
    const float p0 = e0 / s, p1 = e1 / s;
    const int y = labels[m];
    const float inv = 1.0f / (float)M;
    const float g0 = (p0 - (y == 0 ? 1.f : 0.f)) * inv, g1 = (p1 - (y == 1 ? 1.f : 0.f)) * inv;
    dlog[2 * m] = g0;
    dlog[2 * m + 1] = g1;
    dlogb[2 * m] = f2bf(g0);
    dlogb[2 * m + 1] = f2bf(g1);
    loss[m] = -((y == 0 ? l0 : l1) - mx - logf(s));
}

// column sums of G (M x N, f32) -> out[n]; one block per column strip
__global__ void colsum_kernel(const float *G, int64_t M, int64_t N, float *out) {
    const int64_t n = (int64_t)blockIdx.x;
    __shared__ float part[256];
    float s = 0.f;
    for (int64_t m = threadIdx.x; m < M; m += blockDim.x) s += G[m * N + n];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = part[0];
}

// Adam on a float32 master copy; refreshes the bf16 copy used by the products
__global__ void adam_kernel(float *W, const float *G, float *mom, float *vel, int64_t count,
                            float lr, float b1, float b2, float eps, float bc1, float bc2,
                            unsigned short *Wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float g = G[i];
    const float mm = b1 * mom[i] + (1.f - b1) * g;
    const float vv = b2 * vel[i] + (1.f - b2) * g * g;
    mom[i] = mm;
    vel[i] = vv;
    const float w = W[i] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    W[i] = w;
    if (Wb) Wb[i] = f2bf(w);
}

// (theta, omega, T) rows -> bf16 rows of 16:
//   [sin(theta), cos(theta), (omega - c1) * s1, (T - c2) * s2, (theta - c0) * s0,
//    (omega^2 / 2 - cos(theta)) * s1^2, 0, ...]
// The attractor of an IC is 2 pi-periodic in theta, so the network sees theta
// through its sine and cosine (plus the scaled angle itself); the last
// feature is the undamped pendulum energy, whose decay orders the basin bands.
__global__ void prep_input_kernel(const float *x, int64_t M, float c0, float s0, float c1, float s1,
                                  float c2, float s2, unsigned short *out) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * 16) return;
    const int64_t m = idx >> 4;
    const int col = (int)(idx & 15);
    const float th = x[3 * m];
    float v = 0.f;
    if (col == 0) v = sinf(th);
    if (col == 1) v = cosf(th);
    if (col == 2) v = (x[3 * m + 1] - c1) * s1;
    if (col == 3) v = (x[3 * m + 2] - c2) * s2;
    if (col == 4) v = 0.0f * (th - c0) * s0;  // raw angle unused (2 pi symmetry)
    if (col == 5) {
        const float om = x[3 * m + 1];
        v = (0.5f * om * om - cosf(th)) * (s1 * s1);
    }
    out[idx] = f2bf(v);
}

// logits (M x 2 f32) + bias -> label = argmax, confidence = max softmax probability
__global__ void classify_kernel(const float *Z, const float *b, int64_t M, int32_t *label,
                                float *conf) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const float l0 = Z[2 * m] + b[0], l1 = Z[2 * m + 1] + b[1];
    const float d = l1 - l0;
    const float p1 = 1.0f / (1.0f + __expf(-d));
    label[m] = d > 0.f ? 1 : 0;
    conf[m] = fmaxf(p1, 1.0f - p1);
}

}  // namespace mlp
}  // namespace bh

namespace bh {
namespace mlp {

// ---- fused inference -------------------------------------------------------
// One workgroup (4 waves) takes a tile of 128 rows through the whole network:
// input features -> 256 -> 256 -> 256 -> 2 -> (label, confidence).  The
// tile's activations live in one LDS buffer (bf16, padded rows, 66 KB, so
// two workgroups share a CU); a layer reads the whole buffer, waits at a
// barrier and writes its output over it.  Products use the 32-deep bf16
// matrix instruction with the weights as the first operand, so each lane
// ends with four adjacent output columns of one row (one 8-byte LDS store).
// Weights are pre-packed per layer as [kstep][colblock][lane][8] with
//   k = 32 kstep + 8 (lane / 16) + e,  col = 16 colblock + lane % 16,
// one 16-byte load per lane, coalesced across the wave and served from L2.
constexpr int IT_ROWS = 128;
constexpr int IT_LD = 264;  // LDS row stride in bf16 (256 + 8)

typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));

struct InferArgs {
    const float *x;              // (M, 3) theta, omega, T
    int64_t M;
    const bf16x8v *w1, *w2, *w3, *w4;  // packed fragments
    const float *b1, *b2, *b3, *b4;    // biases (b4 has 2 entries)
    float c0, s0, c1, s1, c2, s2;      // input scaling
    int32_t *label;
    float *conf;
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
// SiLU with a hardware reciprocal (1 ulp); the result is rounded to bf16
__device__ __forceinline__ float silu_fast(float z) {
    return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

// out^T blocks of one 256-wide layer: wave w owns column blocks 4w .. 4w+3
// for all 8 row blocks; acc[rb][c][r] = column 16 (4w + c) + 4 (lane / 16) + r
// of row 16 rb + lane % 16.
template <int KSTEPS>
__device__ __forceinline__ void layer_products(const unsigned short *buf, const bf16x8v *W,
                                               int wave, int lane, floatx4 (&acc)[8][4]) {
#pragma unroll
    for (int rb = 0; rb < 8; rb++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[rb][c] = (floatx4){0.f, 0.f, 0.f, 0.f};
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    bf16x8v wf[4];
#pragma unroll
    for (int c = 0; c < 4; c++) wf[c] = W[(4 * wave + c) * 64 + lane];
#pragma unroll 1
    for (int s = 0; s < KSTEPS; s++) {
        bf16x8v wn[4];
        if (s + 1 < KSTEPS) {  // next step's fragments in flight during this step
#pragma unroll
            for (int c = 0; c < 4; c++) wn[c] = W[((int64_t)(s + 1) * 16 + 4 * wave + c) * 64 + lane];
        }
#pragma unroll
        for (int rb = 0; rb < 8; rb++) {
            const bf16x8v xf =
                *reinterpret_cast<const bf16x8v *>(buf + (rb * 16 + lr) * IT_LD + s * 32 + lk);
#pragma unroll
            for (int c = 0; c < 4; c++)
                acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[c], xf, acc[rb][c], 0, 0, 0);
        }
        if (s + 1 < KSTEPS) {
#pragma unroll
            for (int c = 0; c < 4; c++) wf[c] = wn[c];
        }
    }
}

__device__ __forceinline__ void layer_store(unsigned short *buf, const float *b, int wave,
                                            int lane, const floatx4 (&acc)[8][4]) {
    const int lr = lane & 15, q = lane >> 4;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int col = (4 * wave + c) * 16 + 4 * q;
        const float4 bias = *reinterpret_cast<const float4 *>(b + col);
#pragma unroll
        for (int rb = 0; rb < 8; rb++) {
            bf16x4v v;
            v[0] = (__bf16)silu_fast(acc[rb][c][0] + bias.x);
            v[1] = (__bf16)silu_fast(acc[rb][c][1] + bias.y);
            v[2] = (__bf16)silu_fast(acc[rb][c][2] + bias.z);
            v[3] = (__bf16)silu_fast(acc[rb][c][3] + bias.w);
            *reinterpret_cast<bf16x4v *>(buf + (rb * 16 + lr) * IT_LD + col) = v;
        }
    }
}

template <int KSTEPS>
__device__ __forceinline__ void dense_layer(unsigned short *buf, const bf16x8v *W, const float *b,
                                            int wave, int lane) {
    floatx4 acc[8][4];
    layer_products<KSTEPS>(buf, W, wave, lane, acc);
    __syncthreads();  // every wave has read the layer input
    layer_store(buf, b, wave, lane, acc);
    __syncthreads();
}

__global__ void __launch_bounds__(256, 2) mlp_infer_kernel(InferArgs a) {
    __shared__ __attribute__((aligned(16))) unsigned short buf[IT_ROWS * IT_LD];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t row0 = (int64_t)blockIdx.x * IT_ROWS;

    // input features of one row per thread: columns 0..31 (6 used, rest 0)
    if (tid < IT_ROWS) {
        const int64_t m = row0 + tid;
        float f[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (m < a.M) {
            const float th = a.x[3 * m], om = a.x[3 * m + 1], T = a.x[3 * m + 2];
            float sn, cs;
            sincosf(th, &sn, &cs);
            f[0] = sn;
            f[1] = cs;
            f[2] = (om - a.c1) * a.s1;
            f[3] = (T - a.c2) * a.s2;
            f[4] = 0.0f * (th - a.c0) * a.s0;
            f[5] = (0.5f * om * om - cs) * (a.s1 * a.s1);
        }
        bf16x8v v0, z;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            v0[e] = (__bf16)(e < 6 ? f[e] : 0.f);
            z[e] = (__bf16)0.f;
        }
        bf16x8v *dst = reinterpret_cast<bf16x8v *>(buf + tid * IT_LD);
        dst[0] = v0;
        dst[1] = z;
        dst[2] = z;
        dst[3] = z;
    }
    __syncthreads();
    dense_layer<1>(buf, a.w1, a.b1, wave, lane);
    dense_layer<8>(buf, a.w2, a.b2, wave, lane);
    dense_layer<8>(buf, a.w3, a.b3, wave, lane);
    // output layer: 2 logits in a 16-column block; wave w takes row blocks
    // 2w, 2w+1; lanes 0..15 hold logits 0 and 1 of row 16 rb + lane
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    floatx4 acc[2] = {(floatx4){0.f, 0.f, 0.f, 0.f}, (floatx4){0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const bf16x8v wf = a.w4[s * 64 + lane];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int rb = 2 * wave + j;
            const bf16x8v xf = *reinterpret_cast<const bf16x8v *>(buf + (rb * 16 + lr) * IT_LD + s * 32 + lk);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, xf, acc[j], 0, 0, 0);
        }
    }
    if (lane < 16) {
        const float b0 = a.b4[0], b1 = a.b4[1];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int64_t m = row0 + (2 * wave + j) * 16 + lane;
            if (m < a.M) {
                const float d = (acc[j][1] + b1) - (acc[j][0] + b0);
                const float p1 = 1.0f / (1.0f + __expf(-d));
                a.label[m] = d > 0.f ? 1 : 0;
                a.conf[m] = fmaxf(p1, 1.0f - p1);
            }
        }
    }
}

}  // namespace mlp
}  // namespace bh
