// mlp.hip -- learned warm-start of the parquet iteration: a multilayer perceptron that maps
// couplings and Λ to compressed vertex coefficients, trained and evaluated by the kernels
// here, and the FP32 decoder from coefficients to an action.  No reference counterpart: the
// reference starts launch_parquet! from the bare action (Launcher.jl:676-678, parquet.jl:54).
//
// Layers: z_l = h_{l-1} W_l^T + b_l, h_l = silu(z_l), linear output.  Products run on MFMA
// with BF16 operands (weights, activations and back-propagated gradients rounded to BF16,
// round to nearest even) and FP32 accumulation; master weights, biases, Adam moments and the
// pre-activations stay FP32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pffrg.h"

struct pffrg_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<int> dims;     // layer widths: n_in, width, ..., n_out
    std::vector<long> woff;    // offset of W_l (out x in, row-major) in params
    std::vector<long> boff;    // offset of b_l
    long n_params = 0;
    float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t *wbf = nullptr;   // BF16 copy of params (weights used)
    int cap = 0;               // batch capacity of the activation buffers
    std::vector<float *> z;    // pre-activations per layer [cap x dims[l+1]]
    std::vector<uint16_t *> hbf;  // BF16 activations per layer input [cap x dims[l]] (hbf[0] = input)
    float *dz = nullptr, *dh = nullptr, *y = nullptr, *t = nullptr, *loss = nullptr;
    uint16_t *dzbf = nullptr;
    std::string err;
};

struct pffrg_decoder {
    int device = 0;
    hipStream_t stream = nullptr;
    long length = 0;
    int ncoef = 0;
    float *mean = nullptr, *basis = nullptr, *coef = nullptr;
    double *out = nullptr;
    std::string err;
};

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite values)
    return (uint16_t)(u >> 16);
}

// C(m, n) (+)= Σ_k A(m, k) B(k, n) (+ bias[n]); A(m, k) = a[m*sam + k*sak], B(k, n) = b[k*sbk + n*sbn]
// (BF16 bits), C(m, n) = c[m*scm + n*scn] (FP32); out-of-range operands read as zero.
struct GemmArgs {
    const uint16_t *a;
    long sam, sak;
    const uint16_t *b;
    long sbk, sbn;
    int M, N, K;
    float *c;
    long scm, scn;
    const float *bias;
    int accumulate;
};

// One 32 x 32 tile of C per 256-thread block.  K goes in chunks of 128 staged through LDS (A as
// [m][k] and B as [n][k], k contiguous, rows padded to 136); the four waves split a chunk's K
// into 32-deep slices, each wave forming the whole tile's partial product on its slice with
// 2 x 2 v_mfma_f32_16x16x32_bf16 (lane l: A row l%16 and B column l%16 at k = 8 (l/16) .. +7;
// result rows 4 (l/16) .. +3 of column l%16).  The next chunk is loaded into registers while the
// current one is multiplied, and the waves' partial tiles are summed in wave order at the end.
// The small GEMMs of the warm-start network (batch 64, widths up to 1024) fill more CUs this way
// than with 64-wide tiles, and each wave's K loop is a quarter as long.
constexpr int GT = 32, GKC = 128, GLD = GKC + 8;

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// Stage one operand chunk: rows r of the tile (m for A, n for B) and k of the chunk.  kcont: the
// operand is contiguous along k (stride sk == 1); then thread t takes row t/8 and 16 consecutive
// k, else k = t/2 and 16 consecutive rows.  Element (row, k) is p[row*sr + k*sk].
__device__ __forceinline__ void gemm_load(const uint16_t *__restrict__ p, long sr, long sk, int r0, int R, int k0, int K,
                                          bool kcont, int t, uint4 (&v)[2]) {
    const int a0 = kcont ? (t >> 3) : 16 * (t & 1);   // first row
    const int b0 = kcont ? 16 * (t & 7) : (t >> 1);   // first k
    const int row = r0 + a0, k = k0 + b0;
    const long step = kcont ? sk : sr;                 // stride along the 16 elements
    const uint16_t *q = p + (long)row * sr + (long)k * sk;
    const bool full = kcont ? (row < R && k + 15 < K) : (k < K && row + 15 < R);
    if (full && step == 1 && ((reinterpret_cast<uintptr_t>(q) & 15u) == 0)) {
        v[0] = reinterpret_cast<const uint4 *>(q)[0];
        v[1] = reinterpret_cast<const uint4 *>(q)[1];
        return;
    }
    uint16_t e[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int rr = kcont ? row : row + i, kk = kcont ? k + i : k;
        e[i] = (rr < R && kk < K) ? p[(long)rr * sr + (long)kk * sk] : (uint16_t)0;
    }
    uint32_t w[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = (uint32_t)e[2 * i] | ((uint32_t)e[2 * i + 1] << 16);
    v[0] = uint4{w[0], w[1], w[2], w[3]};
    v[1] = uint4{w[4], w[5], w[6], w[7]};
}

// the staged elements into the LDS tile s ([row][k], row stride GLD)
__device__ __forceinline__ void gemm_store(uint16_t *s, bool kcont, int t, const uint4 (&v)[2]) {
    if (kcont) {
        uint4 *d = reinterpret_cast<uint4 *>(s + (t >> 3) * GLD + 16 * (t & 7));
        d[0] = v[0];
        d[1] = v[1];
        return;
    }
    const int a0 = 16 * (t & 1), k = t >> 1;
    const uint32_t w[8] = {v[0].x, v[0].y, v[0].z, v[0].w, v[1].x, v[1].y, v[1].z, v[1].w};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        s[(a0 + 2 * i) * GLD + k] = (uint16_t)(w[i] & 0xffffu);
        s[(a0 + 2 * i + 1) * GLD + k] = (uint16_t)(w[i] >> 16);
    }
}

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
    __shared__ __align__(16) uint16_t sA[GT * GLD];
    __shared__ __align__(16) uint16_t sB[GT * GLD];
    __shared__ float sR[4 * GT * GT];
    const int t = threadIdx.x, wv = t >> 6, l = t & 63;
    const int tm = blockIdx.y * GT, tn = blockIdx.x * GT;
    const bool akc = g.sak == 1, bkc = g.sbk == 1;
    v4f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    uint4 ra[2], rb[2];
    gemm_load(g.a, g.sam, g.sak, tm, g.M, 0, g.K, akc, t, ra);
    gemm_load(g.b, g.sbn, g.sbk, tn, g.N, 0, g.K, bkc, t, rb);
    for (int k0 = 0; k0 < g.K; k0 += GKC) {
        __syncthreads();  // the previous chunk's products have read the tiles
        gemm_store(sA, akc, t, ra);
        gemm_store(sB, bkc, t, rb);
        __syncthreads();
        if (k0 + GKC < g.K) {
            gemm_load(g.a, g.sam, g.sak, tm, g.M, k0 + GKC, g.K, akc, t, ra);
            gemm_load(g.b, g.sbn, g.sbk, tn, g.N, k0 + GKC, g.K, bkc, t, rb);
        }
        const int ko = 32 * wv + 8 * (l >> 4);
        bf8 av[2], bv[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            av[i] = *reinterpret_cast<const bf8 *>(sA + (16 * i + (l & 15)) * GLD + ko);
            bv[i] = *reinterpret_cast<const bf8 *>(sB + (16 * i + (l & 15)) * GLD + ko);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    // the waves' partial tiles, then their sum in wave order
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                sR[wv * GT * GT + (16 * i + 4 * (l >> 4) + q) * GT + 16 * j + (l & 15)] = acc[i][j][q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < GT * GT / 256; ++q) {
        const int e = t + 256 * q, mr = e / GT, nc = e % GT;
        const int m = tm + mr, n = tn + nc;
        if (m < g.M && n < g.N) {
            float v = ((sR[e] + sR[GT * GT + e]) + sR[2 * GT * GT + e]) + sR[3 * GT * GT + e];
            if (g.bias) v += g.bias[n];
            float *p = g.c + m * g.scm + n * g.scn;
            *p = g.accumulate ? *p + v : v;
        }
    }
}

__device__ inline float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

// h = silu(z) as BF16 for the next product
__global__ void k_silu_bf(const float *z, uint16_t *h, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        h[i] = f2bf(z[i] * sigm(z[i]));
}

// dz = dh * silu'(z), and its BF16 copy
__global__ void k_silu_back(const float *dh, const float *z, float *dz, uint16_t *dzbf, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float s = sigm(z[i]);
        const float d = dh[i] * (s * (1.0f + z[i] * (1.0f - s)));
        dz[i] = d;
        dzbf[i] = f2bf(d);
    }
}

__global__ void k_to_bf(const float *x, uint16_t *y, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

// mean-squared error L = Σ (y - t)^2 / n and dL/dy = 2 (y - t) / n: each block's partial sum of
// squares (a fixed grid for a given n, reduced in a fixed order) into part[blockIdx.x], then
// k_mse_sum adds the partials in block order, so the loss does not depend on scheduling
constexpr int MSE_BLOCKS = 64;

__global__ __launch_bounds__(256) void k_mse(const float *y, const float *t, long n, float *dy, uint16_t *dybf,
                                             float *part) {
    __shared__ float red[256];
    float s = 0.f;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const float d = y[i] - t[i];
        s += d * d;
        const float g = 2.0f * d / (float)n;
        dy[i] = g;
        dybf[i] = f2bf(g);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(64) void k_mse_sum(const float *part, int np, long n, float *loss) {
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int b = 0; b < np; ++b) s += part[b];
        *loss = s / (float)n;
    }
}

// db[n] = Σ_b dz[b][n]: 32 columns per 256-thread block, eight row slices per column summed in
// slice order
__global__ __launch_bounds__(256) void k_colsum(const float *dz, int B, int N, float *db) {
    __shared__ float red[8][33];
    const int c = threadIdx.x & 31, r = threadIdx.x >> 5;
    const int n = blockIdx.x * 32 + c;
    float s = 0.f;
    if (n < N)
        for (int b = r; b < B; b += 8) s += dz[(long)b * N + n];
    red[r][c] = s;
    __syncthreads();
    if (r == 0 && n < N) {
        float v = red[0][c];
#pragma unroll
        for (int k = 1; k < 8; ++k) v += red[k][c];
        db[n] = v;
    }
}

__global__ void k_adam(float *p, const float *g, float *m1, float *m2, uint16_t *pbf, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.0f - b1) * gi;
        const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        const float q = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
        p[i] = q;
        pbf[i] = f2bf(q);
    }
}

// out[j] = mean[j] + Σ_k coef[k] basis[k][j]  (FP32 products, FP64 output)
__global__ void k_decode(const float *__restrict__ mean, const float *__restrict__ basis, const float *__restrict__ coef,
                         int ncoef, long length, double *out) {
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < length; j += (long)gridDim.x * blockDim.x) {
        float s = mean[j];
        for (int k = 0; k < ncoef; ++k) s += coef[k] * basis[(long)k * length + j];
        out[j] = (double)s;
    }
}

int grid1(long n) {
    long b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

int mfail(pffrg_mlp *m, int code, const std::string &s) {
    if (m) m->err = s;
    return code;
}

#define MCHK(m, expr)                                                                                     \
    do {                                                                                                  \
        hipError_t _e = (expr);                                                                           \
        if (_e != hipSuccess) return mfail((m), PFFRG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

hipError_t gemm(hipStream_t s, const uint16_t *a, long sam, long sak, const uint16_t *b, long sbk, long sbn, int M, int N,
                int K, float *c, long scm, long scn, const float *bias, int accumulate) {
    GemmArgs g{a, sam, sak, b, sbk, sbn, M, N, K, c, scm, scn, bias, accumulate};
    hipLaunchKernelGGL(k_gemm_bf16, dim3((N + GT - 1) / GT, (M + GT - 1) / GT), dim3(256), 0, s, g);
    return hipGetLastError();
}

int ensure_batch(pffrg_mlp *m, int B) {
    if (B <= m->cap) return 0;
    for (float *p : m->z) (void)hipFree(p);
    for (uint16_t *p : m->hbf) (void)hipFree(p);
    (void)hipFree(m->dz);
    (void)hipFree(m->dh);
    (void)hipFree(m->y);
    (void)hipFree(m->t);
    (void)hipFree(m->dzbf);
    const int L = (int)m->dims.size() - 1;
    int wmax = 0;
    for (int d : m->dims) wmax = d > wmax ? d : wmax;
    m->z.assign(L, nullptr);
    m->hbf.assign(L, nullptr);
    for (int l = 0; l < L; ++l) {
        MCHK(m, hipMalloc(&m->z[l], (size_t)B * m->dims[l + 1] * sizeof(float)));
        MCHK(m, hipMalloc(&m->hbf[l], (size_t)B * m->dims[l] * sizeof(uint16_t)));
    }
    MCHK(m, hipMalloc(&m->dz, (size_t)B * wmax * sizeof(float)));
    MCHK(m, hipMalloc(&m->dh, (size_t)B * wmax * sizeof(float)));
    MCHK(m, hipMalloc(&m->dzbf, (size_t)B * wmax * sizeof(uint16_t)));
    MCHK(m, hipMalloc(&m->y, (size_t)B * m->dims[L] * sizeof(float)));
    MCHK(m, hipMalloc(&m->t, (size_t)B * m->dims[L] * sizeof(float)));
    m->cap = B;
    return 0;
}

// forward pass of a batch already in hbf[0]; output pre-activations of the last layer in z[L-1]
int forward_dev(pffrg_mlp *m, int B) {
    const int L = (int)m->dims.size() - 1;
    for (int l = 0; l < L; ++l) {
        const int in = m->dims[l], out = m->dims[l + 1];
        MCHK(m, gemm(m->stream, m->hbf[l], in, 1, m->wbf + m->woff[l], 1, in, B, out, in, m->z[l], out, 1,
                     m->params + m->boff[l], 0));
        if (l + 1 < L) {
            hipLaunchKernelGGL(k_silu_bf, dim3(grid1((long)B * out)), dim3(256), 0, m->stream, m->z[l], m->hbf[l + 1],
                               (long)B * out);
            MCHK(m, hipGetLastError());
        }
    }
    return 0;
}

int upload_input(pffrg_mlp *m, const float *x, int B) {
    float *tmp = m->dh;  // staging (cap x wmax >= B x n_in)
    MCHK(m, hipMemcpyAsync(tmp, x, (size_t)B * m->dims[0] * sizeof(float), hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_to_bf, dim3(grid1((long)B * m->dims[0])), dim3(256), 0, m->stream, tmp, m->hbf[0],
                       (long)B * m->dims[0]);
    MCHK(m, hipGetLastError());
    return 0;
}

// Box-Muller normals from splitmix64 (host, for the initial weights)
uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int pffrg_mlp_create(int32_t n_in, int32_t width, int32_t n_layers, int32_t n_out, uint64_t seed, int32_t device,
                     pffrg_mlp **out) {
    if (!out || n_in < 1 || width < 1 || n_layers < 2 || n_out < 1 || width > 8192 || n_out > 65536) return PFFRG_ERR_ARG;
    *out = nullptr;
    pffrg_mlp *m = new pffrg_mlp();
    m->device = device;
    m->dims.push_back(n_in);
    for (int l = 0; l + 1 < n_layers; ++l) m->dims.push_back(width);
    m->dims.push_back(n_out);
    long off = 0;
    for (size_t l = 0; l + 1 < m->dims.size(); ++l) {
        m->woff.push_back(off);
        off += (long)m->dims[l] * m->dims[l + 1];
        m->boff.push_back(off);
        off += m->dims[l + 1];
    }
    m->n_params = off;
    // He-normal weights (std sqrt(2 / fan_in)), the output layer scaled by 0.1; zero biases
    std::vector<float> p(off, 0.0f);
    uint64_t s = seed;
    for (size_t l = 0; l + 1 < m->dims.size(); ++l) {
        const double sd = std::sqrt(2.0 / m->dims[l]) * (l + 2 == m->dims.size() ? 0.1 : 1.0);
        for (long i = 0; i < (long)m->dims[l] * m->dims[l + 1]; i += 2) {
            const double u1 = (double)(splitmix(s) >> 11) * 0x1.0p-53, u2 = (double)(splitmix(s) >> 11) * 0x1.0p-53;
            const double r = std::sqrt(-2.0 * std::log1p(-u1));
            p[m->woff[l] + i] = (float)(sd * r * std::cos(2.0 * M_PI * u2));
            if (i + 1 < (long)m->dims[l] * m->dims[l + 1]) p[m->woff[l] + i + 1] = (float)(sd * r * std::sin(2.0 * M_PI * u2));
        }
    }
    std::vector<uint16_t> pb(off);
    for (long i = 0; i < off; ++i) pb[i] = f2bf(p[i]);
    bool ok = hipSetDevice(device) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&m->params, off * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&m->grads, off * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&m->m1, off * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&m->m2, off * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&m->wbf, off * sizeof(uint16_t)) == hipSuccess;
    ok = ok && hipMalloc(&m->loss, (1 + MSE_BLOCKS) * sizeof(float)) == hipSuccess;  // the loss, then k_mse's partials
    ok = ok && hipMemcpy(m->params, p.data(), off * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(m->wbf, pb.data(), off * sizeof(uint16_t), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemset(m->grads, 0, off * sizeof(float)) == hipSuccess;
    ok = ok && hipMemset(m->m1, 0, off * sizeof(float)) == hipSuccess;
    ok = ok && hipMemset(m->m2, 0, off * sizeof(float)) == hipSuccess;
    if (!ok) {
        pffrg_mlp_destroy(m);
        return PFFRG_ERR_HIP;
    }
    *out = m;
    return 0;
}

void pffrg_mlp_destroy(pffrg_mlp *m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    for (float *p : m->z) (void)hipFree(p);
    for (uint16_t *p : m->hbf) (void)hipFree(p);
    for (void *p : {(void *)m->params, (void *)m->grads, (void *)m->m1, (void *)m->m2, (void *)m->wbf, (void *)m->dz,
                    (void *)m->dh, (void *)m->y, (void *)m->t, (void *)m->dzbf, (void *)m->loss})
        (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char *pffrg_mlp_last_error(const pffrg_mlp *m) { return m ? m->err.c_str() : "null mlp"; }

int64_t pffrg_mlp_num_params(const pffrg_mlp *m) { return m ? m->n_params : -1; }

int pffrg_mlp_get_params(pffrg_mlp *m, float *params) {
    if (!m || !params) return PFFRG_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipMemcpy(params, m->params, m->n_params * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int pffrg_mlp_set_params(pffrg_mlp *m, const float *params) {
    if (!m || !params) return PFFRG_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    std::vector<uint16_t> pb(m->n_params);
    for (long i = 0; i < m->n_params; ++i) pb[i] = f2bf(params[i]);
    MCHK(m, hipMemcpy(m->params, params, m->n_params * sizeof(float), hipMemcpyHostToDevice));
    MCHK(m, hipMemcpy(m->wbf, pb.data(), m->n_params * sizeof(uint16_t), hipMemcpyHostToDevice));
    MCHK(m, hipMemset(m->m1, 0, m->n_params * sizeof(float)));
    MCHK(m, hipMemset(m->m2, 0, m->n_params * sizeof(float)));
    return 0;
}

int pffrg_mlp_forward(pffrg_mlp *m, const float *x, int32_t batch, float *y) {
    if (!m || !x || !y || batch < 1) return PFFRG_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    int rc;
    if ((rc = ensure_batch(m, batch)) || (rc = upload_input(m, x, batch)) || (rc = forward_dev(m, batch))) return rc;
    const int L = (int)m->dims.size() - 1;
    MCHK(m, hipMemcpyAsync(y, m->z[L - 1], (size_t)batch * m->dims[L] * sizeof(float), hipMemcpyDeviceToHost,
                           m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return 0;
}

int pffrg_mlp_backward(pffrg_mlp *m, const float *x, const float *target, int32_t batch, float *loss, float *grads) {
    if (!m || !x || !target || batch < 1) return PFFRG_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    int rc;
    if ((rc = ensure_batch(m, batch)) || (rc = upload_input(m, x, batch)) || (rc = forward_dev(m, batch))) return rc;
    const int L = (int)m->dims.size() - 1;
    const int B = batch;
    const long no = (long)B * m->dims[L];
    MCHK(m, hipMemcpyAsync(m->t, target, no * sizeof(float), hipMemcpyHostToDevice, m->stream));
    const int nb = (int)std::min<long>(MSE_BLOCKS, (no + 255) / 256);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, m->stream, m->z[L - 1], m->t, no, m->dz, m->dzbf, m->loss + 1);
    hipLaunchKernelGGL(k_mse_sum, dim3(1), dim3(64), 0, m->stream, m->loss + 1, nb, no, m->loss);
    MCHK(m, hipGetLastError());
    for (int l = L - 1; l >= 0; --l) {
        const int in = m->dims[l], out = m->dims[l + 1];
        // dW_l[n][k] = Σ_b dz[b][n] h_{l-1}[b][k]
        MCHK(m, gemm(m->stream, m->dzbf, 1, out, m->hbf[l], in, 1, out, in, B, m->grads + m->woff[l], in, 1, nullptr, 0));
        hipLaunchKernelGGL(k_colsum, dim3((out + 31) / 32), dim3(256), 0, m->stream, m->dz, B, out,
                           m->grads + m->boff[l]);
        MCHK(m, hipGetLastError());
        if (l > 0) {
            // dh[b][k] = Σ_n dz[b][n] W[n][k], then dz_{l-1} = dh silu'(z_{l-1})
            MCHK(m, gemm(m->stream, m->dzbf, out, 1, m->wbf + m->woff[l], in, 1, B, in, out, m->dh, in, 1, nullptr, 0));
            hipLaunchKernelGGL(k_silu_back, dim3(grid1((long)B * in)), dim3(256), 0, m->stream, m->dh, m->z[l - 1], m->dz,
                               m->dzbf, (long)B * in);
            MCHK(m, hipGetLastError());
        }
    }
    if (loss) MCHK(m, hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    if (grads) MCHK(m, hipMemcpyAsync(grads, m->grads, m->n_params * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return 0;
}

int pffrg_mlp_adam(pffrg_mlp *m, float lr, float beta1, float beta2, float eps, int32_t step) {
    if (!m || step < 1) return PFFRG_ERR_ARG;
    MCHK(m, hipSetDevice(m->device));
    const float c1 = 1.0f - std::pow(beta1, (float)step), c2 = 1.0f - std::pow(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(grid1(m->n_params)), dim3(256), 0, m->stream, m->params, m->grads, m->m1, m->m2,
                       m->wbf, m->n_params, lr, beta1, beta2, eps, c1, c2);
    MCHK(m, hipGetLastError());
    MCHK(m, hipStreamSynchronize(m->stream));
    return 0;
}

int pffrg_decoder_create(int64_t length, int32_t n_coef, const float *mean, const float *basis, int32_t device,
                         pffrg_decoder **out) {
    if (!out || !mean || !basis || length < 1 || n_coef < 1) return PFFRG_ERR_ARG;
    *out = nullptr;
    pffrg_decoder *d = new pffrg_decoder();
    d->device = device;
    d->length = length;
    d->ncoef = n_coef;
    bool ok = hipSetDevice(device) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&d->mean, length * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&d->basis, (size_t)length * n_coef * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&d->coef, n_coef * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&d->out, length * sizeof(double)) == hipSuccess;
    ok = ok && hipMemcpy(d->mean, mean, length * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(d->basis, basis, (size_t)length * n_coef * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        pffrg_decoder_destroy(d);
        return PFFRG_ERR_HIP;
    }
    *out = d;
    return 0;
}

void pffrg_decoder_destroy(pffrg_decoder *d) {
    if (!d) return;
    (void)hipSetDevice(d->device);
    (void)hipFree(d->mean);
    (void)hipFree(d->basis);
    (void)hipFree(d->coef);
    (void)hipFree(d->out);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

int pffrg_decoder_apply(pffrg_decoder *d, const float *coef, double *out) {
    if (!d || !coef || !out) return PFFRG_ERR_ARG;
    if (hipSetDevice(d->device) != hipSuccess) return PFFRG_ERR_HIP;
    if (hipMemcpyAsync(d->coef, coef, d->ncoef * sizeof(float), hipMemcpyHostToDevice, d->stream) != hipSuccess)
        return PFFRG_ERR_HIP;
    hipLaunchKernelGGL(k_decode, dim3(grid1(d->length)), dim3(256), 0, d->stream, d->mean, d->basis, d->coef, d->ncoef,
                       d->length, d->out);
    if (hipGetLastError() != hipSuccess) return PFFRG_ERR_HIP;
    if (hipMemcpyAsync(out, d->out, d->length * sizeof(double), hipMemcpyDeviceToHost, d->stream) != hipSuccess)
        return PFFRG_ERR_HIP;
    return hipStreamSynchronize(d->stream) == hipSuccess ? 0 : PFFRG_ERR_HIP;
}

}  // extern "C"
