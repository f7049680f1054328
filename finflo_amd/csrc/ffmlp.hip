// The learned EARSM closure: a 5 -> 64 -> 64 -> 3 SiLU network that maps the invariants of
// the T-scaled strain and rotation tensors to EARSM1's coefficients (beta4, beta6, beta9),
// from which N = -beta6/beta9 replaces the cubic solve and N correction of earsm.f:170-193.  Cells whose normalised
// invariants fall outside the box spanned by the training samples take the exact EARSM.
//
// Every product of the network runs on the matrix cores: v_mfma_f32_16x16x32_bf16 for the
// contractions over hidden features, v_mfma_f32_16x16x16_bf16 for the rest, BF16 operands and
// FP32 accumulation, FP32 biases and activations, FP32 master weights.  The
// network is evaluated transposed (features along M, cells along N), so a layer's 16x16
// accumulator block is, lane for lane, the B operand of the next layer:
//   A operand  a[i] = A[m = lane % 16][k = 4 (lane / 16) + i]
//   B operand  b[i] = B[k = 4 (lane / 16) + i][n = lane % 16]
//   result     c[i] = C[m = 4 (lane / 16) + i][n = lane % 16]
// The weight gradients contract over cells; their operands are the stored activations and
// signals read back transposed from LDS.  One wave works on 16 cells at a time.
//
// This file is part of the ffhip.hip translation unit (included at its end).

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int MNIN = 5, MNH = 64, MNOUT = 3;
constexpr int MOFF_W1 = 0, MOFF_B1 = MOFF_W1 + MNH * MNIN, MOFF_W2 = MOFF_B1 + MNH,
              MOFF_B2 = MOFF_W2 + MNH * MNH, MOFF_W3 = MOFF_B2 + MNH,
              MOFF_B3 = MOFF_W3 + MNOUT * MNH, MNP = MOFF_B3 + MNOUT;
static_assert(MNP == FF_MLP_NPARAM, "parameter count of the header");
constexpr int MPP = MNP + 1;           // per-wave partial: gradient, then the loss
// normalisation block: x mean (5), x scale (5), y mean (3), y scale (3), hull low (5), high (5)
constexpr int MN_XM = 0, MN_XS = 5, MN_YM = 10, MN_YS = 13, MN_LO = 16, MN_HI = 21;
static_assert(MN_HI + 5 == FF_MLP_NNORM, "normalisation block of the header");

// round to nearest even (gfx950's conversion instruction)
__device__ __forceinline__ short bf16_bits(float f)
{
    return __builtin_bit_cast(short, (__bf16)f);
}

__device__ __forceinline__ float bf16_val(short h)
{
    return __uint_as_float(((unsigned)(unsigned short)h) << 16);
}

// SiLU z s(z) and its derivative s (1 + z (1 - s)), s = 1/(1 + e^-z), with the hardware's
// exp2 and reciprocal (the activations are rounded to BF16 next)
__device__ __forceinline__ float sigm_f(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_f(float z) { return z * sigm_f(z); }

__device__ __forceinline__ float dsilu_f(float z)
{
    const float s = sigm_f(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// 16x16x32: a[j] = A[m = lane % 16][k = 8 (lane / 16) + j], b[j] = B[k = 8 (lane / 16) + j]
// [n = lane % 16], j = 0..7.  The hidden layers contract over 64 features in two such steps.
// Their B operand is two accumulator blocks of the previous layer side by side, so the
// instruction's k = 8g + j stands for feature phi(kb, g, j) = 32 kb + 16 (j / 4) + 4 g + j % 4;
// the weights are gathered in the same order and no lane exchanges values.
__device__ __forceinline__ v4f mfma32(v8s a, v8s b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                   __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

__device__ __forceinline__ int phi(int kb, int g, int j) { return 32 * kb + 16 * (j >> 2) + 4 * g + (j & 3); }

__device__ __forceinline__ v8s cat8(v4s lo, v4s hi)
{
    return v8s{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// The weights of one wave as A operands (forward: W1, W2, W3; backward: W2^T, W3^T)
struct MlpFrags {
    v4s w1[4], w3t[4];                        // 16x16x16 operands
    v8s w2[4][2], w3[2], w2t[4][2];           // 16x16x32 operands, k in phi order
    float b1[4][4], b2[4][4], b3[4];          // this lane's rows of the biases
};

__device__ __forceinline__ void mlp_load_frags(const float *p, int lane, MlpFrags &f, bool bwd)
{
    const int g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int i = 0; i < 4; i++) f.b3[i] = 4 * g + i < MNOUT ? p[MOFF_B3 + 4 * g + i] : 0.f;
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = 4 * g + i;
            f.b1[ob][i] = p[MOFF_B1 + 16 * ob + k];
            f.b2[ob][i] = p[MOFF_B2 + 16 * ob + k];
            f.w1[ob][i] = bf16_bits(k < MNIN ? p[MOFF_W1 + (16 * ob + c) * MNIN + k] : 0.f);
            if (bwd) f.w3t[ob][i] = bf16_bits(k < MNOUT ? p[MOFF_W3 + k * MNH + 16 * ob + c] : 0.f);
        }
#pragma unroll
        for (int kb = 0; kb < 2; kb++)
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = phi(kb, g, j);
                f.w2[ob][kb][j] = bf16_bits(p[MOFF_W2 + (16 * ob + c) * MNH + k]);
                if (bwd) f.w2t[ob][kb][j] = bf16_bits(p[MOFF_W2 + k * MNH + 16 * ob + c]);
            }
    }
#pragma unroll
    for (int kb = 0; kb < 2; kb++)
#pragma unroll
        for (int j = 0; j < 8; j++)
            f.w3[kb][j] = bf16_bits(c < MNOUT ? p[MOFF_W3 + c * MNH + phi(kb, g, j)] : 0.f);
}

// Forward pass of 16 cells.  xn: this lane's normalised inputs 4g..4g+3 of cell lane%16.
// Returns z1, z2 (pre-activations) and the BF16 activations, and the normalised outputs y
// (rows 4g+i < 3 of lanes with g = 0).
struct MlpFwd {
    v4f z1[4], z2[4], y;
    v4s x, h1[4], h2[4];
};

__device__ __forceinline__ void mlp_forward16(const MlpFrags &f, int lane, const float xn[4],
                                              MlpFwd &o)
{
    const int g = lane >> 4;
#pragma unroll
    for (int i = 0; i < 4; i++) o.x[i] = bf16_bits(4 * g + i < MNIN ? xn[i] : 0.f);
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
        v4f z = {0.f, 0.f, 0.f, 0.f};
        z = mfma(f.w1[ob], o.x, z);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            z[i] += f.b1[ob][i];
            o.h1[ob][i] = bf16_bits(silu_f(z[i]));
        }
        o.z1[ob] = z;
    }
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
        v4f z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 2; kb++) z = mfma32(f.w2[ob][kb], cat8(o.h1[2 * kb], o.h1[2 * kb + 1]), z);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            z[i] += f.b2[ob][i];
            o.h2[ob][i] = bf16_bits(silu_f(z[i]));
        }
        o.z2[ob] = z;
    }
    v4f y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) y = mfma32(f.w3[kb], cat8(o.h2[2 * kb], o.h2[2 * kb + 1]), y);
#pragma unroll
    for (int i = 0; i < 4; i++) y[i] += f.b3[i];
    o.y = y;
}

// Feature and target transforms: the invariants span several decades with either sign, so
// the network sees asinh of each; the coefficients have fixed signs (beta4, beta6 < 0,
// beta9 > 0) and span many decades, so it learns log(sign * beta).
__device__ __forceinline__ float mlp_feature(float v)
{
    const float a = fabsf(v);
    return copysignf(__logf(a + __fsqrt_rn(a * a + 1.f)), v);       // asinh
}
__device__ __forceinline__ float mlp_sign(int q) { return q < 2 ? -1.f : 1.f; }
__device__ __forceinline__ float mlp_target(float beta, int q) { return logf(fmaxf(mlp_sign(q) * beta, 1e-37f)); }
__device__ __forceinline__ float mlp_output(float yn, const float *nrm, int q)
{
    return mlp_sign(q) * expf(yn * nrm[MN_YS + q] + nrm[MN_YM + q]);
}

// this lane's normalised inputs 4g..4g+3 of cell `cell` (zero beyond the five)
__device__ __forceinline__ void mlp_inputs(const float *x, const float *nrm, int cell, bool ok,
                                           int g, float xn[4])
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int k = 4 * g + i;
        xn[i] = (ok && k < MNIN) ? (mlp_feature(x[(size_t)cell * MNIN + k]) - nrm[MN_XM + k]) / nrm[MN_XS + k] : 0.f;
    }
}

// Inference on samples: physical outputs (beta4, beta6, beta9) of n cells.  Wave w takes
// cells 16 w .. 16 w + 15 in steps of the grid.
__global__ void __launch_bounds__(256) k_mlp_infer(const float *p, const float *nrm, const float *x,
                                                   float *y, int n)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    MlpFrags f;
    mlp_load_frags(p, lane, f, false);
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int nw = gridDim.x * (blockDim.x >> 6);
    for (int base = 16 * wave; base < n; base += 16 * nw) {
        const int cell = base + c;
        float xn[4];
        mlp_inputs(x, nrm, cell, cell < n, g, xn);
        MlpFwd o;
        mlp_forward16(f, lane, xn, o);
        if (g == 0 && cell < n)
#pragma unroll
            for (int i = 0; i < MNOUT; i++) y[(size_t)cell * MNOUT + i] = mlp_output(o.y[i], nrm, i);
    }
}

// Per-wave LDS for the transposes of the weight gradients: [feature][cell] in BF16
struct __attribute__((aligned(16))) MlpLds {
    short x[16][16], h1[64][16], h2[64][16], dz1[64][16], dz2[64][16], dy[16][16];
};

__device__ __forceinline__ void lds_put(short (*m)[16], int row0, int lane, v4s v)
{
    const int g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int i = 0; i < 4; i++) m[row0 + 4 * g + i][c] = v[i];
}

// A or B operand read back transposed: rows row0 + lane%16, cells 4g..4g+3
__device__ __forceinline__ v4s lds_get(short (*m)[16], int row0, int lane)
{
    const int g = lane >> 4, c = lane & 15;
    return *(const v4s *)&m[row0 + c][4 * g];
}

__device__ __forceinline__ float xor_sum16(float v)
{
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Loss and gradient partials over a set of samples.  Sample s of the set is sample
// idx[s] when idx is given, else s.  The loss is the mean over `ntot` samples of the squared
// error of the normalised outputs summed over the outputs.  Each wave writes its partial
// gradient and loss to part[wave * MPP ..].
__global__ void __launch_bounds__(256) k_mlp_grad(const float *p, const float *nrm, const float *x,
                                                  const float *t, const unsigned *idx, int nset,
                                                  float inv_ntot, float *part)
{
    __shared__ MlpLds lds_all[4];
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    MlpLds &L = lds_all[threadIdx.x >> 6];
    MlpFrags f;
    mlp_load_frags(p, lane, f, true);
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int nw = gridDim.x * (blockDim.x >> 6);
    v4f gw1[4], gw2[4][4], gw3[4];
    float gb1[4][4], gb2[4][4], gb3[4], loss = 0.f;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        gw1[a] = v4f{0.f, 0.f, 0.f, 0.f};
        gw3[a] = v4f{0.f, 0.f, 0.f, 0.f};
        gb3[a] = 0.f;
#pragma unroll
        for (int b2 = 0; b2 < 4; b2++) {
            gw2[a][b2] = v4f{0.f, 0.f, 0.f, 0.f};
            gb1[a][b2] = 0.f;
            gb2[a][b2] = 0.f;
        }
    }
    for (int base = 16 * wave; base < nset; base += 16 * nw) {
        const int s = base + c;
        const bool ok = s < nset;
        const int cell = ok ? (idx ? (int)idx[s] : s) : 0;
        float xn[4];
        mlp_inputs(x, nrm, cell, ok, g, xn);
        MlpFwd o;
        mlp_forward16(f, lane, xn, o);
        // output signal dL/dyn = 2 (yn - tn) / N on rows o < 3 (lanes g = 0)
        v4f dy = {0.f, 0.f, 0.f, 0.f};
        v4s dyb;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int q = 4 * g + i;
            if (ok && q < MNOUT) {
                const float tn = (mlp_target(t[(size_t)cell * MNOUT + q], q) - nrm[MN_YM + q]) / nrm[MN_YS + q];
                const float r = o.y[i] - tn;
                loss += r * r;
                dy[i] = 2.f * r * inv_ntot;
            }
            dyb[i] = bf16_bits(dy[i]);
            gb3[i] += dy[i];
        }
        // back through layer 3 and the second activation
        v4s dz2b[4];
#pragma unroll
        for (int fb = 0; fb < 4; fb++) {
            v4f dh = {0.f, 0.f, 0.f, 0.f};
            dh = mfma(f.w3t[fb], dyb, dh);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float dz = dh[i] * dsilu_f(o.z2[fb][i]);
                gb2[fb][i] += dz;
                dz2b[fb][i] = bf16_bits(dz);
            }
        }
        // back through layer 2 and the first activation
        v4s dz1b[4];
#pragma unroll
        for (int fb = 0; fb < 4; fb++) {
            v4f dh = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < 2; kb++) dh = mfma32(f.w2t[fb][kb], cat8(dz2b[2 * kb], dz2b[2 * kb + 1]), dh);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float dz = dh[i] * dsilu_f(o.z1[fb][i]);
                gb1[fb][i] += dz;
                dz1b[fb][i] = bf16_bits(dz);
            }
        }
        // weight gradients: contractions over the 16 cells, operands transposed through LDS
        lds_put(L.x, 0, lane, o.x);
        lds_put(L.dy, 0, lane, dyb);
#pragma unroll
        for (int b4 = 0; b4 < 4; b4++) {
            lds_put(L.h1, 16 * b4, lane, o.h1[b4]);
            lds_put(L.h2, 16 * b4, lane, o.h2[b4]);
            lds_put(L.dz1, 16 * b4, lane, dz1b[b4]);
            lds_put(L.dz2, 16 * b4, lane, dz2b[b4]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        const v4s xt = lds_get(L.x, 0, lane), dyt = lds_get(L.dy, 0, lane);
#pragma unroll
        for (int ob = 0; ob < 4; ob++) {
            const v4s d2 = lds_get(L.dz2, 16 * ob, lane), d1 = lds_get(L.dz1, 16 * ob, lane);
            gw1[ob] = mfma(d1, xt, gw1[ob]);
#pragma unroll
            for (int ib = 0; ib < 4; ib++) gw2[ob][ib] = mfma(d2, lds_get(L.h1, 16 * ib, lane), gw2[ob][ib]);
            gw3[ob] = mfma(dyt, lds_get(L.h2, 16 * ob, lane), gw3[ob]);
        }
        // the LDS tiles are rewritten by the next 16 cells only after these reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
    // partials: weight blocks lane by lane, biases summed over the 16 cell lanes
    float *o = part + (size_t)wave * MPP;
#pragma unroll
    for (int ob = 0; ob < 4; ob++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = 16 * ob + 4 * g + i;
            if (c < MNIN) o[MOFF_W1 + row * MNIN + c] = gw1[ob][i];
#pragma unroll
            for (int ib = 0; ib < 4; ib++) o[MOFF_W2 + row * MNH + 16 * ib + c] = gw2[ob][ib][i];
            if (4 * g + i < MNOUT) o[MOFF_W3 + (4 * g + i) * MNH + 16 * ob + c] = gw3[ob][i];
            const float s1 = xor_sum16(gb1[ob][i]), s2 = xor_sum16(gb2[ob][i]);
            if (c == 0) {
                o[MOFF_B1 + row] = s1;
                o[MOFF_B2 + row] = s2;
            }
        }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float s3 = xor_sum16(gb3[i]);
        if (c == 0 && 4 * g + i < MNOUT) o[MOFF_B3 + 4 * g + i] = s3;
    }
    float l = loss;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) l += __shfl_xor(l, m, 64);
    if (lane == 0) o[MNP] = l * inv_ntot;
}

// Sum of the wave partials in a fixed order; with lr > 0 one Adam step on the parameters
// (beta1 0.9, beta2 0.999, eps 1e-8; bias-corrected with step number `step` >= 1)
__global__ void __launch_bounds__(256) k_mlp_reduce(float *p, float *m, float *v, const float *part,
                                                    int nwave, float lr, int step, float *grad)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= MPP) return;
    float s = 0.f;
    for (int w = 0; w < nwave; w++) s += part[(size_t)w * MPP + j];
    if (grad) grad[j] = s;
    if (lr > 0.f && j < MNP) {
        const float b1 = 0.9f, b2 = 0.999f;
        const float mj = b1 * m[j] + (1.f - b1) * s;
        const float vj = b2 * v[j] + (1.f - b2) * s * s;
        m[j] = mj;
        v[j] = vj;
        const float mh = mj / (1.f - powf(b1, (float)step)), vh = vj / (1.f - powf(b2, (float)step));
        p[j] -= lr * mh / (sqrtf(vh) + 1e-8f);
    }
}

// Minibatch indices: a hash of (seed, step, slot), uniform over [0, n)
__global__ void __launch_bounds__(256) k_mlp_batch(unsigned *idx, int nb, int n, unsigned seed, int step)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nb) return;
    unsigned long long h = ((unsigned long long)seed << 32) ^ ((unsigned long long)step * 0x9E3779B97F4A7C15ull) ^ (unsigned long long)s;
    h ^= h >> 33; h *= 0xFF51AFD7ED558CCDull; h ^= h >> 33; h *= 0xC4CEB9FE1A85EC53ull; h ^= h >> 33;
    idx[s] = (unsigned)(h % (unsigned long long)n);
}

// Normalisation block from samples: means and standard deviations of the transformed inputs
// and outputs, then the box of the normalised inputs.  Three grid-wide passes, each a fixed
// split of the samples over NRM_PARTS workgroups and a fixed-order tree inside each,
// followed by a one-workgroup pass over the partials (FP64 sums: deterministic, independent
// of the launch timing).
constexpr int NRM_COLS = MNIN + MNOUT;
constexpr int NRM_PARTS = 256;                        // workgroups of a partial pass
constexpr int NRM_SCR = 2 * NRM_COLS * NRM_PARTS + NRM_COLS;

__device__ __forceinline__ double nrm_value(const float *x, const float *y, int r, int col)
{
    return col < MNIN ? (double)mlp_feature(x[(size_t)r * MNIN + col])
                      : (double)mlp_target(y[(size_t)r * MNOUT + col - MNIN], col - MNIN);
}

// fixed-order tree over the workgroup's 256 values (op 0 sum, 1 min, 2 max)
template <int OP>
__device__ __forceinline__ double nrm_tree(double *red, double v)
{
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] = OP == 0 ? red[t] + red[t + h] : OP == 1 ? fmin(red[t], red[t + h])
                                                                    : fmax(red[t], red[t + h]);
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// PASS 0: sums of every column; 1: squared deviations from the means (scr tail); 2: min
// and max of the normalised inputs (nrm's means and deviations, as the network uses them)
template <int PASS>
__global__ void __launch_bounds__(256) k_mlp_norm_part(const float *x, const float *y, int n,
                                                       const float *nrm, double *scr)
{
    __shared__ double red[256];
    const int t = threadIdx.x;
    const int r0 = (int)((long long)n * blockIdx.x / NRM_PARTS);
    const int r1 = (int)((long long)n * (blockIdx.x + 1) / NRM_PARTS);
    const double *mean = scr + 2 * NRM_COLS * NRM_PARTS;
    for (int col = 0; col < (PASS < 2 ? NRM_COLS : MNIN); col++) {
        double a = PASS < 2 ? 0. : 1e300, z = -1e300;
        for (int r = r0 + t; r < r1; r += 256) {
            if (PASS == 0) {
                a += nrm_value(x, y, r, col);
            } else if (PASS == 1) {
                const double d = nrm_value(x, y, r, col) - mean[col];
                a += d * d;
            } else {
                const float v = (mlp_feature(x[(size_t)r * MNIN + col]) - nrm[MN_XM + col]) / nrm[MN_XS + col];
                a = fmin(a, (double)v);
                z = fmax(z, (double)v);
            }
        }
        if (PASS < 2) {
            a = nrm_tree<0>(red, a);
            if (t == 0) scr[col * NRM_PARTS + blockIdx.x] = a;
        } else {
            a = nrm_tree<1>(red, a);
            z = nrm_tree<2>(red, z);
            if (t == 0) {
                scr[(2 * col) * NRM_PARTS + blockIdx.x] = a;
                scr[(2 * col + 1) * NRM_PARTS + blockIdx.x] = z;
            }
        }
    }
}

// one workgroup: the partials of a pass, in a fixed order, into the block
template <int PASS>
__global__ void __launch_bounds__(256) k_mlp_norm_fin(int n, float *nrm, double *scr)
{
    __shared__ double red[256];
    static_assert(NRM_PARTS == 256, "one partial per thread");
    const int t = threadIdx.x;
    double *mean = scr + 2 * NRM_COLS * NRM_PARTS;
    for (int col = 0; col < (PASS < 2 ? NRM_COLS : MNIN); col++) {
        const int q = col < MNIN ? col : col - MNIN;
        if (PASS < 2) {
            const double sum = nrm_tree<0>(red, scr[col * NRM_PARTS + t]);
            if (t == 0) {
                if (PASS == 0) {
                    mean[col] = sum / n;
                    nrm[col < MNIN ? MN_XM + q : MN_YM + q] = (float)(sum / n);
                } else {
                    double sd = sqrt(sum / n);
                    if (!(sd > 0.)) sd = 1.;
                    nrm[col < MNIN ? MN_XS + q : MN_YS + q] = (float)sd;
                }
            }
        } else {
            const double lo = nrm_tree<1>(red, scr[(2 * col) * NRM_PARTS + t]);
            const double hi = nrm_tree<2>(red, scr[(2 * col + 1) * NRM_PARTS + t]);
            if (t == 0) {
                nrm[MN_LO + q] = (float)lo;
                nrm[MN_HI + q] = (float)hi;
            }
        }
    }
}

__device__ __forceinline__ void earsm_inputs(const EarsmT &e, float v[MNIN])
{
    v[0] = (float)e.sii; v[1] = (float)e.wii; v[2] = (float)e.iiis;
    v[3] = (float)e.swwiv; v[4] = (float)e.sswwv;
}

// Samples of the exact closure on the interior cells of a block: inputs (the five
// invariants) and targets (beta4, beta6, beta9), cell c = (K-1) JMAX IMAX + (J-1) IMAX + I-1
__global__ void __launch_bounds__(256) k_earsm_samples(Blk b, float *x, float *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.imax) return;
    const int L = cid(b, i, j, k);
    const size_t cell = ((size_t)(k - 1) * b.jmax + (j - 1)) * b.imax + (i - 1);
    EarsmT e;
    earsm_terms(b, L, e);
    double beta[4];
    earsm_betas(e, beta);
    float v[MNIN];
    earsm_inputs(e, v);
    for (int q = 0; q < MNIN; q++) x[cell * MNIN + q] = v[q];
    for (int q = 0; q < MNOUT; q++) y[cell * MNOUT + q] = (float)beta[q + 1];
}

// EARSM1 with the learned closure on the interior cells: the wave evaluates the network on
// its 64 cells (16 at a time).  A cell whose normalised invariants lie in the training box
// takes N = -beta6/beta9 from the network's coefficients in place of the cubic solve and N
// correction, and the four betas from N by EARSM1's own formulas (earsm.f:196-205), so that
// they stay consistent with each other (b_ij is a sum of cancelling terms); any other cell
// takes the exact EARSM1.  Cells are numbered I fastest over the interior.
__global__ void __launch_bounds__(256) k_earsm_mlp(Blk b, const float *p, const float *nrm)
{
    __shared__ float xs[4][64][MNIN];
    __shared__ float ys[4][64][MNOUT];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
    const int ncell = b.imax * b.jmax * b.kmax;
    const int cell = (blockIdx.x * 4 + w) * 64 + lane;
    const bool ok = cell < ncell;
    int L = 0;
    EarsmT e;
    bool in = false;
    if (ok) {
        const int i = cell % b.imax + 1, j = (cell / b.imax) % b.jmax + 1, k = cell / (b.imax * b.jmax) + 1;
        L = cid(b, i, j, k);
        earsm_terms(b, L, e);
        float v[MNIN];
        earsm_inputs(e, v);
        in = true;
        for (int q = 0; q < MNIN; q++) {
            const float xn = (mlp_feature(v[q]) - nrm[MN_XM + q]) / nrm[MN_XS + q];
            in = in && xn >= nrm[MN_LO + q] && xn <= nrm[MN_HI + q];
            xs[w][lane][q] = xn;
        }
    } else {
        for (int q = 0; q < MNIN; q++) xs[w][lane][q] = 0.f;
    }
    __syncthreads();
    MlpFrags f;
    mlp_load_frags(p, lane, f, false);
#pragma unroll 1
    for (int cb = 0; cb < 4; cb++) {
        float xn[4];
#pragma unroll
        for (int i = 0; i < 4; i++) xn[i] = 4 * g + i < MNIN ? xs[w][16 * cb + c][4 * g + i] : 0.f;
        MlpFwd o;
        mlp_forward16(f, lane, xn, o);
        if (g == 0)
#pragma unroll
            for (int i = 0; i < MNOUT; i++) ys[w][16 * cb + c][i] = mlp_output(o.y[i], nrm, i);
    }
    __syncthreads();
    if (!ok) return;
    double beta[4];
    const double b6 = ys[w][lane][1], b9 = ys[w][lane][2];
    if (in && b6 < 0. && b9 > 0.)
        earsm_betas_n(e, -b6 / b9, beta);
    else
        earsm_betas(e, beta);
    earsm_store(b, L, e, beta);
}

static void mlp_free(MlpRec &r)
{
    for (void *q : {(void *)r.p, (void *)r.m, (void *)r.v, (void *)r.nrm, (void *)r.nscr,
                    (void *)r.part, (void *)r.idx, (void *)r.grad})
        if (q) (void)hipFree(q);
    r = MlpRec{};
}

static int mlp_alloc(ff_ctx *c, MlpRec &r)
{
    if (r.p) return FF_OK;
    if (hipMalloc(&r.p, sizeof(float) * MNP) != hipSuccess ||
        hipMalloc(&r.m, sizeof(float) * MNP) != hipSuccess ||
        hipMalloc(&r.v, sizeof(float) * MNP) != hipSuccess ||
        hipMalloc(&r.nrm, sizeof(float) * FF_MLP_NNORM) != hipSuccess ||
        hipMalloc(&r.grad, sizeof(float) * MPP) != hipSuccess) {
        mlp_free(r);
        return fail(c, FF_E_NOMEM, "learned closure parameters%s");
    }
    HIPCHK(c, hipMemsetAsync(r.m, 0, sizeof(float) * MNP, c->stream));
    HIPCHK(c, hipMemsetAsync(r.v, 0, sizeof(float) * MNP, c->stream));
    return FF_OK;
}

static int mlp_workspace(ff_ctx *c, MlpRec &r, int waves, int nidx)
{
    if (waves > r.part_waves) {
        if (r.part) (void)hipFree(r.part);
        r.part = nullptr;
        r.part_waves = 0;
        if (hipMalloc(&r.part, sizeof(float) * (size_t)MPP * waves) != hipSuccess)
            return fail(c, FF_E_NOMEM, "learned closure gradient partials%s");
        r.part_waves = waves;
    }
    if (nidx > r.idx_cap) {
        if (r.idx) (void)hipFree(r.idx);
        r.idx = nullptr;
        r.idx_cap = 0;
        if (hipMalloc(&r.idx, sizeof(unsigned) * nidx) != hipSuccess)
            return fail(c, FF_E_NOMEM, "learned closure minibatch%s");
        r.idx_cap = nidx;
    }
    return FF_OK;
}

// gradient over `nset` samples (all of them, or the minibatch in r.idx), reduced into r.grad;
// with lr > 0 an Adam step as well
static int mlp_grad_step(ff_ctx *c, MlpRec &r, const float *x, const float *y, const unsigned *idx,
                         int nset, float lr)
{
    const int blocks = std::min(std::max((nset + 255) / 256, 1), 512);
    const int waves = blocks * 4;
    int rc = mlp_workspace(c, r, waves, 0);
    if (rc) return rc;
    if ((rc = launch(c, "k_mlp_grad", k_mlp_grad, dim3(blocks), dim3(256), (const float *)r.p,
                     (const float *)r.nrm, x, y, idx, nset, 1.0f / (float)nset, r.part))) return rc;
    if (lr > 0.f) r.step++;
    return launch(c, "k_mlp_reduce", k_mlp_reduce, dim3((MPP + 255) / 256), dim3(256), r.p, r.m, r.v,
                  (const float *)r.part, waves, lr, r.step, r.grad);
}

extern "C" {

int ff_mlp_set(ff_ctx *c, const float *params, const float *norm)
{
    if (!c || (!params && !norm)) return FF_E_ARG;
    int rc = mlp_alloc(c, c->mlp);
    if (rc) return rc;
    if (params) {
        HIPCHK(c, hipMemcpyAsync(c->mlp.p, params, sizeof(float) * MNP, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemsetAsync(c->mlp.m, 0, sizeof(float) * MNP, c->stream));
        HIPCHK(c, hipMemsetAsync(c->mlp.v, 0, sizeof(float) * MNP, c->stream));
        c->mlp.step = 0;
    }
    if (norm) {
        HIPCHK(c, hipMemcpyAsync(c->mlp.nrm, norm, sizeof(float) * FF_MLP_NNORM, hipMemcpyHostToDevice, c->stream));
        c->mlp.ready = true;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FF_OK;
}

int ff_mlp_get(ff_ctx *c, float *params, float *norm)
{
    if (!c) return FF_E_ARG;
    if (!c->mlp.p) return fail(c, FF_E_STATE, "learned closure not set%s");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (params) HIPCHK(c, hipMemcpy(params, c->mlp.p, sizeof(float) * MNP, hipMemcpyDeviceToHost));
    if (norm) HIPCHK(c, hipMemcpy(norm, c->mlp.nrm, sizeof(float) * FF_MLP_NNORM, hipMemcpyDeviceToHost));
    return FF_OK;
}

int ff_earsm_samples(ff_ctx *c, int level, int block, float *x, float *y)
{
    if (!c || !x || !y) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    if (c->prm.iturb != 6 || c->prm.istres != 1) return fail(c, FF_E_UNSUP, "EARSM samples need ITURB = 6, ISTRES = 1%s");
    c->lev = level;
    const Blk &b = r->dev;
    int rc = launch(c, "k_earsm_samples", k_earsm_samples, cells_grid(b, 256), dim3(256), b, x, y);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FF_OK;
}

int ff_mlp_fit_norm(ff_ctx *c, const float *x, const float *y, int n)
{
    if (!c || !x || !y || n < 1) return FF_E_ARG;
    int rc = mlp_alloc(c, c->mlp);
    if (rc) return rc;
    MlpRec &m = c->mlp;
    if (!m.nscr && hipMalloc(&m.nscr, sizeof(double) * NRM_SCR) != hipSuccess)
        return fail(c, FF_E_NOMEM, "normalisation partials%s");
    const dim3 gp(NRM_PARTS), t256(256);
    if ((rc = launch(c, "k_mlp_norm", k_mlp_norm_part<0>, gp, t256, x, y, n, (const float *)m.nrm, m.nscr)) ||
        (rc = launch(c, "k_mlp_norm", k_mlp_norm_fin<0>, dim3(1), t256, n, m.nrm, m.nscr)) ||
        (rc = launch(c, "k_mlp_norm", k_mlp_norm_part<1>, gp, t256, x, y, n, (const float *)m.nrm, m.nscr)) ||
        (rc = launch(c, "k_mlp_norm", k_mlp_norm_fin<1>, dim3(1), t256, n, m.nrm, m.nscr)) ||
        (rc = launch(c, "k_mlp_norm", k_mlp_norm_part<2>, gp, t256, x, y, n, (const float *)m.nrm, m.nscr)) ||
        (rc = launch(c, "k_mlp_norm", k_mlp_norm_fin<2>, dim3(1), t256, n, m.nrm, m.nscr)))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->mlp.ready = true;
    return FF_OK;
}

int ff_mlp_infer(ff_ctx *c, const float *x, int n, float *y)
{
    if (!c || !x || !y || n < 0) return FF_E_ARG;
    if (!c->mlp.ready) return fail(c, FF_E_STATE, "learned closure not set%s");
    const int blocks = std::min(std::max((n + 63) / 64, 1), 4096);
    int rc = launch(c, "k_mlp_infer", k_mlp_infer, dim3(blocks), dim3(256), (const float *)c->mlp.p,
                    (const float *)c->mlp.nrm, x, y, n);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FF_OK;
}

int ff_mlp_grad(ff_ctx *c, const float *x, const float *y, int n, float *grad)
{
    if (!c || !x || !y || n < 1 || !grad) return FF_E_ARG;
    if (!c->mlp.ready) return fail(c, FF_E_STATE, "learned closure not set%s");
    int rc = mlp_grad_step(c, c->mlp, x, y, nullptr, n, 0.f);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(grad, c->mlp.grad, sizeof(float) * MPP, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FF_OK;
}

int ff_mlp_train(ff_ctx *c, const float *x, const float *y, int n, int steps, int batch, float lr,
                 unsigned seed, float *loss)
{
    if (!c || !x || !y || n < 1 || steps < 0 || batch < 16 || !(lr > 0.f)) return FF_E_ARG;
    if (!c->mlp.ready) return fail(c, FF_E_STATE, "learned closure not set%s");
    int rc = mlp_workspace(c, c->mlp, 0, batch);
    if (rc) return rc;
    for (int s = 0; s < steps; s++) {
        if ((rc = launch(c, "k_mlp_batch", k_mlp_batch, dim3((batch + 255) / 256), dim3(256), c->mlp.idx,
                         batch, n, seed, s))) return rc;
        if ((rc = mlp_grad_step(c, c->mlp, x, y, c->mlp.idx, batch, lr))) return rc;
        if (loss)
            HIPCHK(c, hipMemcpyAsync(loss + s, c->mlp.grad + MNP, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FF_OK;
}

}  // extern "C"

// EARSM1 through the learned closure (called by ff_turb after k_turb_sst skipped EARSM1)
static int earsm_mlp_pass(ff_ctx *c, const Blk &b)
{
    const int ncell = b.imax * b.jmax * b.kmax;
    return launch(c, "k_earsm_mlp", k_earsm_mlp, dim3((ncell + 255) / 256), dim3(256), b,
                  (const float *)c->mlp.p, (const float *)c->mlp.nrm);
}
