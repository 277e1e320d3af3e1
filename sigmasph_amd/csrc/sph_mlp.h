// sph_mlp.h -- the constitutive surrogate's network, 12 -> 64 -> 64 -> 7 with ReLU, on
// the bf16 matrix cores (included once, at the end of sigmasph.hip).
//
// One wave evaluates 16 samples with v_mfma_f32_16x16x16_bf16: operands are bf16, sums
// f32.  Operand layouts of the 16x16x16 instruction (lane l, r = l % 16, q = l / 16):
//   A (M x K): lane holds A[r][4q .. 4q+3]       B (K x N): lane holds B[4q .. 4q+3][r]
//   C (M x N): lane holds C[4q .. 4q+3][r]
// so a layer's output (samples x units, C layout) is already the B operand of the next
// weight-gradient product and, transposed, its A operand; the forward products need the
// A layout of the activations and go through LDS once per layer.
//
// Parameters (f32 master, packed): W1[64][16] (inputs padded 12 -> 16 with zero columns),
// b1[64], W2[64][64], b2[64], W3[16][64] (outputs padded 7 -> 16 with zero rows),
// b3[16]; a bf16 copy of the three matrices feeds the matrix cores.  Inputs are
// normalised as (x - mean) / std and outputs de-normalised as y std + mean with the
// vectors in sph_mlp.norm; the loss is the mean square error in the normalised space.

#define MLP_IN 16
#define MLP_H 64
#define MLP_OUT 16
#define MLP_NIN 12
#define MLP_NOUT 7
#define MLP_W1 0
#define MLP_B1 (MLP_W1 + MLP_H * MLP_IN)
#define MLP_W2 (MLP_B1 + MLP_H)
#define MLP_B2 (MLP_W2 + MLP_H * MLP_H)
#define MLP_W3 (MLP_B2 + MLP_H)
#define MLP_B3 (MLP_W3 + MLP_OUT * MLP_H)
#define MLP_NPARAM (MLP_B3 + MLP_OUT)
// bf16 copies: W1, W2, W3 back to back
#define MLP_BW1 0
#define MLP_BW2 (MLP_BW1 + MLP_H * MLP_IN)
#define MLP_BW3 (MLP_BW2 + MLP_H * MLP_H)
#define MLP_NBF (MLP_BW3 + MLP_OUT * MLP_H)
// normalisation vectors: in mean [16], in 1/std [16], out mean [16], out std [16]
#define MLP_NORM_IN_MEAN 0
#define MLP_NORM_IN_ISTD 16
#define MLP_NORM_OUT_MEAN 32
#define MLP_NORM_OUT_STD 48
#define MLP_NNORM 64

typedef short mlp_bf4 __attribute__((__vector_size__(4 * sizeof(short))));
typedef float mlp_f4 __attribute__((__vector_size__(4 * sizeof(float))));

// round-to-nearest-even float -> bf16 bits
__device__ __forceinline__ unsigned short mlp_bf16(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ mlp_bf4 mlp_pack(float a, float b, float c, float d) {
    mlp_bf4 v;
    v[0] = (short)mlp_bf16(a); v[1] = (short)mlp_bf16(b); v[2] = (short)mlp_bf16(c); v[3] = (short)mlp_bf16(d);
    return v;
}
__device__ __forceinline__ mlp_bf4 mlp_load4(const unsigned short* p) {
    mlp_bf4 v;
    v[0] = (short)p[0]; v[1] = (short)p[1]; v[2] = (short)p[2]; v[3] = (short)p[3];
    return v;
}
__device__ __forceinline__ mlp_f4 mlp_mfma(mlp_bf4 a, mlp_bf4 b, mlp_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mlp_f4 mlp_zero4() {
    mlp_f4 z = {0.f, 0.f, 0.f, 0.f};
    return z;
}

struct MlpDev {
    const float* w;               // f32 master (biases read from here)
    const unsigned short* wb;     // bf16 matrices
    const float* norm;
};

// The forward pass of one wave over samples s0 .. s0+15 of a row-major [n][16] input
// (already normalised when `normed`); leaves the layer outputs in C layout for the
// backward pass.  `lds` is this wave's [16][64] f32 scratch.
struct MlpWave {
    mlp_f4 h1[4], h2[4], y;   // C layout: [sample 4q+i][unit 16t + r]
    int r, q;

    __device__ void forward(const MlpDev& m, const float* x, int64_t n, int64_t s0, float* lds, bool normed) {
        const int l = threadIdx.x % 64;
        r = l % 16;
        q = l / 16;
        // inputs, A layout: sample s0 + r, features 4q .. 4q+3
        float xa[4];
        const int64_t s = s0 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int f = 4 * q + i;
            float v = 0.f;
            if (s < n && f < MLP_NIN) {
                v = x[s * MLP_IN + f];
                if (!normed) v = (v - m.norm[MLP_NORM_IN_MEAN + f]) * m.norm[MLP_NORM_IN_ISTD + f];
            }
            xa[i] = v;
        }
        const mlp_bf4 a0 = mlp_pack(xa[0], xa[1], xa[2], xa[3]);
        // layer 1: Z1[s][u] = sum_f X[s][f] W1[u][f]; B[k=f][n=u] = W1[16t + r][4q + i]
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            mlp_f4 z = mlp_mfma(a0, mlp_load4(m.wb + MLP_BW1 + (16 * t + r) * MLP_IN + 4 * q), mlp_zero4());
            const float b = m.w[MLP_B1 + 16 * t + r];
#pragma unroll
            for (int i = 0; i < 4; ++i) z[i] = fmaxf(z[i] + b, 0.f);
            h1[t] = z;
        }
        // layer 2 through the A layout of H1
        to_lds(h1, lds);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            mlp_f4 z = mlp_zero4();
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const float* p = lds + r * MLP_H + 16 * kk + 4 * q;
                z = mlp_mfma(mlp_pack(p[0], p[1], p[2], p[3]),
                             mlp_load4(m.wb + MLP_BW2 + (16 * t + r) * MLP_H + 16 * kk + 4 * q), z);
            }
            const float b = m.w[MLP_B2 + 16 * t + r];
#pragma unroll
            for (int i = 0; i < 4; ++i) z[i] = fmaxf(z[i] + b, 0.f);
            h2[t] = z;
        }
        __syncthreads();
        to_lds(h2, lds);
        // layer 3: Y[s][o] = sum_u H2[s][u] W3[o][u]; B[k=u][n=o] = W3[r][16kk + 4q + i]
        mlp_f4 z = mlp_zero4();
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const float* p = lds + r * MLP_H + 16 * kk + 4 * q;
            z = mlp_mfma(mlp_pack(p[0], p[1], p[2], p[3]), mlp_load4(m.wb + MLP_BW3 + r * MLP_H + 16 * kk + 4 * q), z);
        }
        const float b = m.w[MLP_B3 + r];
#pragma unroll
        for (int i = 0; i < 4; ++i) z[i] += b;
        y = z;      // normalised outputs, [sample 4q+i][output r]
        __syncthreads();
    }

    // C-layout tiles [sample][unit] -> lds[sample][unit]; the block synchronises after
    __device__ void to_lds(const mlp_f4* c, float* lds) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) lds[(4 * q + i) * MLP_H + 16 * t + r] = c[t][i];
        __syncthreads();
    }
};

// inference: y[n][8] de-normalised (7 used)
__global__ __launch_bounds__(256) void k_mlp_forward(MlpDev m, const float* __restrict__ x, float* __restrict__ y,
                                                     int64_t n) {
    __shared__ float lds[4][16 * MLP_H];
    const int w = threadIdx.x / 64;
    const int64_t s0 = ((int64_t)blockIdx.x * 4 + w) * 16;
    MlpWave net;
    net.forward(m, x, n, s0, lds[w], false);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t s = s0 + 4 * net.q + i;
        if (s < n && net.r < MLP_NOUT)
            y[s * 8 + net.r] = net.y[i] * m.norm[MLP_NORM_OUT_STD + net.r] + m.norm[MLP_NORM_OUT_MEAN + net.r];
    }
}

// one training pass over n samples: forward, MSE gradient in the normalised space, and
// the parameter gradients summed into grad (per block in LDS, then to global memory);
// loss[0] += sum of squared errors / (n * 7)
__global__ __launch_bounds__(256) void k_mlp_grad(MlpDev m, const float* __restrict__ x, const float* __restrict__ t,
                                                  int64_t n, float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ float lds[4][16 * MLP_H];
    __shared__ float g[MLP_NPARAM];
    __shared__ float lsum[4];
    for (int p = threadIdx.x; p < MLP_NPARAM; p += blockDim.x) g[p] = 0.f;
    const int w = threadIdx.x / 64;
    const int64_t s0 = ((int64_t)blockIdx.x * 4 + w) * 16;
    MlpWave net;
    net.forward(m, x, n, s0, lds[w], false);
    const int r = net.r, q = net.q;
    const float scale = 2.f / (float)(n * MLP_NOUT);
    // dY, C layout [sample 4q+i][output r]
    mlp_f4 dy;
    float se = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t s = s0 + 4 * q + i;
        float d = 0.f;
        if (s < n && r < MLP_NOUT) {
            const float tn = (t[s * 8 + r] - m.norm[MLP_NORM_OUT_MEAN + r]) / m.norm[MLP_NORM_OUT_STD + r];
            d = net.y[i] - tn;
            se += d * d;
        }
        dy[i] = d * scale;
    }
    // layer 3 gradients: dW3[o][u] = sum_s dY[s][o] H2[s][u]: A = dY (C layout), B = H2 (C layout)
    const mlp_bf4 ady = mlp_pack(dy[0], dy[1], dy[2], dy[3]);
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
        mlp_f4 c = mlp_mfma(ady, mlp_pack(net.h2[tt][0], net.h2[tt][1], net.h2[tt][2], net.h2[tt][3]), mlp_zero4());
#pragma unroll
        for (int i = 0; i < 4; ++i) atomicAdd(&g[MLP_W3 + (4 * q + i) * MLP_H + 16 * tt + r], c[i]);
    }
    atomicAdd(&g[MLP_B3 + r], dy[0] + dy[1] + dy[2] + dy[3]);
    // dH2[s][u] = sum_o dY[s][o] W3[o][u]: A = dY in A layout (through LDS), B[k=o][n=u] = W3[4q+i][16tt + r]
    float* L = lds[w];
#pragma unroll
    for (int i = 0; i < 4; ++i) L[(4 * q + i) * 16 + r] = dy[i];
    __syncthreads();
    const mlp_bf4 ady2 = mlp_pack(L[r * 16 + 4 * q], L[r * 16 + 4 * q + 1], L[r * 16 + 4 * q + 2], L[r * 16 + 4 * q + 3]);
    mlp_f4 dz2[4];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
        const unsigned short* wb = m.wb + MLP_BW3;
        mlp_bf4 b;
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = (short)wb[(4 * q + i) * MLP_H + 16 * tt + r];
        mlp_f4 c = mlp_mfma(ady2, b, mlp_zero4());
#pragma unroll
        for (int i = 0; i < 4; ++i) c[i] = net.h2[tt][i] > 0.f ? c[i] : 0.f;
        dz2[tt] = c;
        atomicAdd(&g[MLP_B2 + 16 * tt + r], c[0] + c[1] + c[2] + c[3]);
    }
    // dW2[u2][u1] = sum_s dZ2[s][u2] H1[s][u1]
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const mlp_bf4 a = mlp_pack(dz2[mt][0], dz2[mt][1], dz2[mt][2], dz2[mt][3]);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            mlp_f4 c = mlp_mfma(a, mlp_pack(net.h1[nt][0], net.h1[nt][1], net.h1[nt][2], net.h1[nt][3]), mlp_zero4());
#pragma unroll
            for (int i = 0; i < 4; ++i) atomicAdd(&g[MLP_W2 + (16 * mt + 4 * q + i) * MLP_H + 16 * nt + r], c[i]);
        }
    }
    // dH1[s][u1] = sum_u2 dZ2[s][u2] W2[u2][u1]: A = dZ2 in A layout (LDS), B[k=u2][n=u1] = W2[16kk+4q+i][16nt+r]
    __syncthreads();
    net.to_lds(dz2, L);
    mlp_f4 dz1[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        mlp_f4 c = mlp_zero4();
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const float* p = L + r * MLP_H + 16 * kk + 4 * q;
            mlp_bf4 b;
#pragma unroll
            for (int i = 0; i < 4; ++i) b[i] = (short)m.wb[MLP_BW2 + (16 * kk + 4 * q + i) * MLP_H + 16 * nt + r];
            c = mlp_mfma(mlp_pack(p[0], p[1], p[2], p[3]), b, c);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) c[i] = net.h1[nt][i] > 0.f ? c[i] : 0.f;
        dz1[nt] = c;
        atomicAdd(&g[MLP_B1 + 16 * nt + r], c[0] + c[1] + c[2] + c[3]);
    }
    // dW1[u1][f] = sum_s dZ1[s][u1] X[s][f]: B[k=s][n=f] = X[s0 + 4q + i][r] (normalised)
    mlp_bf4 bx;
    {
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t s = s0 + 4 * q + i;
            v[i] = (s < n && r < MLP_NIN)
                       ? (x[s * MLP_IN + r] - m.norm[MLP_NORM_IN_MEAN + r]) * m.norm[MLP_NORM_IN_ISTD + r]
                       : 0.f;
        }
        bx = mlp_pack(v[0], v[1], v[2], v[3]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        mlp_f4 c = mlp_mfma(mlp_pack(dz1[mt][0], dz1[mt][1], dz1[mt][2], dz1[mt][3]), bx, mlp_zero4());
#pragma unroll
        for (int i = 0; i < 4; ++i) atomicAdd(&g[MLP_W1 + (16 * mt + 4 * q + i) * MLP_IN + r], c[i]);
    }
    // the wave's squared error
    se = (float)wave_sum((double)se);
    if (threadIdx.x % 64 == 0) lsum[w] = se;
    __syncthreads();
    for (int p = threadIdx.x; p < MLP_NPARAM; p += blockDim.x)
        if (g[p] != 0.f) atomicAdd(&grad[p], g[p]);
    if (threadIdx.x == 0) atomicAdd(loss, (lsum[0] + lsum[1] + lsum[2] + lsum[3]) / (float)(n * MLP_NOUT));
}

// Adam on the f32 master, refreshing the bf16 matrices; clears the gradient
__global__ void k_mlp_adam(float* w, unsigned short* wb, float* grad, float* m1, float* m2, float lr, float b1,
                           float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float gr = grad[p];
    grad[p] = 0.f;
    const float a = b1 * m1[p] + (1.f - b1) * gr;
    const float b = b2 * m2[p] + (1.f - b2) * gr * gr;
    m1[p] = a;
    m2[p] = b;
    const float v = w[p] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[p] = v;
    int k = -1;
    if (p < MLP_B1) k = MLP_BW1 + p;
    else if (p >= MLP_W2 && p < MLP_B2) k = MLP_BW2 + (p - MLP_W2);
    else if (p >= MLP_W3 && p < MLP_B3) k = MLP_BW3 + (p - MLP_W3);
    if (k >= 0) wb[k] = mlp_bf16(v);
}

// bf16 copies of the matrices from the master
__global__ void k_mlp_sync(const float* w, unsigned short* wb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    int k = -1;
    if (p < MLP_B1) k = MLP_BW1 + p;
    else if (p >= MLP_W2 && p < MLP_B2) k = MLP_BW2 + (p - MLP_W2);
    else if (p >= MLP_W3 && p < MLP_B3) k = MLP_BW3 + (p - MLP_W3);
    if (k >= 0) wb[k] = mlp_bf16(w[p]);
}

static bool mlp_ok(const sph_mlp* net) {
    return net && net->w && net->wb && net->norm;
}

extern "C" int sph_mlp_sync(const sph_mlp* net, void* stream) {
    if (!mlp_ok(net)) return SPH_EINVAL;
    hipLaunchKernelGGL(k_mlp_sync, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, net->w,
                       (unsigned short*)net->wb);
    return hipGetLastError() == hipSuccess ? SPH_OK : SPH_EHIP;
}

extern "C" int sph_mlp_forward(const sph_mlp* net, const float* x, float* y, int64_t n, void* stream) {
    if (!mlp_ok(net) || n < 0 || (n > 0 && (!x || !y))) return SPH_EINVAL;
    if (n == 0) return SPH_OK;
    MlpDev m{net->w, (const unsigned short*)net->wb, net->norm};
    const int64_t blocks = (n + 63) / 64;
    if (blocks > 0x7fffffff) return SPH_EINVAL;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, m, x, y, n);
    return hipGetLastError() == hipSuccess ? SPH_OK : SPH_EHIP;
}

extern "C" int sph_mlp_train_step(const sph_mlp* net, const float* x, const float* t, int64_t n, float lr,
                                  int32_t step, float* loss, void* stream) {
    if (!mlp_ok(net) || !net->grad || !net->m1 || !net->m2 || !x || !t || !loss || n <= 0 || step < 1)
        return SPH_EINVAL;
    hipStream_t hs = (hipStream_t)stream;
    MlpDev m{net->w, (const unsigned short*)net->wb, net->norm};
    const int64_t blocks = (n + 63) / 64;
    if (blocks > 0x7fffffff) return SPH_EINVAL;
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)blocks), dim3(256), 0, hs, m, x, t, n, net->grad, loss);
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, hs, net->w,
                       (unsigned short*)net->wb, net->grad, net->m1, net->m2, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError() == hipSuccess ? SPH_OK : SPH_EHIP;
}

extern "C" int sph_mlp_gradient(const sph_mlp* net, const float* x, const float* t, int64_t n, float* loss,
                                void* stream) {
    if (!mlp_ok(net) || !net->grad || !x || !t || !loss || n <= 0) return SPH_EINVAL;
    MlpDev m{net->w, (const unsigned short*)net->wb, net->norm};
    const int64_t blocks = (n + 63) / 64;
    if (blocks > 0x7fffffff) return SPH_EINVAL;
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, m, x, t, n, net->grad,
                       loss);
    return hipGetLastError() == hipSuccess ? SPH_OK : SPH_EHIP;
}
