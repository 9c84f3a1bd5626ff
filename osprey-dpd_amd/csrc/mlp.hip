// mlp.hip -- surrogate network: BF16 matrix-core forward/backward and FP32 Adam on gfx950.
//
// Every matrix product runs on v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32 accumulation)
// in LDS-staged workgroup tiles (k_gemm_tile).  Operand fragment layout of that instruction
// for lane l (0..63): A holds A[l%16][4(l/16) .. 4(l/16)+3], B holds
// B[4(l/16) .. 4(l/16)+3][l%16], and the accumulator holds C[4(l/16)+i][l%16], i = 0..3.
// Strided operand addressing gives C = A B, A^T B and A B^T from one kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "dpd_mlp.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));


__host__ __device__ inline uint16_t f2bf(float f)
{
    // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// C[m][n] (ldc) = sum_k A(m,k) B(k,n);  A(m,k) = A[m*am + k*ak], B(k,n) = B[k*bk + n*bn],
// with am or ak equal to 1 and bk or bn equal to 1 (each operand contiguous along one of its
// dimensions).  A workgroup of GM x GN wavefronts computes a TM x TN tile, each wavefront
// FM x FN fragments of 16 x 16.  K advances in chunks of kKC staged in LDS, both operands as
// [row][k] rows (k contiguous, padded by 4 to spread the banks): the global loads are 8-byte
// vectors along each operand's contiguous dimension (coalesced across the workgroup), and
// each fragment operand is one 8-byte LDS read.  blockIdx.z takes the K range
// [z kper, (z + 1) kper); with several, the tile goes to partial buffer z (C + z cstride) and
// k_sum_partials adds them in order.  M, N, K, kper: multiples of 16.
constexpr int kKC = 32;
constexpr int kLDK = kKC + 4;
template <int FM, int FN, int GM, int GN>
__global__ __launch_bounds__(64 * GM * GN) void k_gemm_tile(int M, int N, int K, int kper,
                                                            const uint16_t* __restrict__ A, int am, int ak,
                                                            const uint16_t* __restrict__ B, int bk, int bn,
                                                            float* __restrict__ C, int ldc, int64_t cstride)
{
    constexpr int TM = 16 * FM * GM, TN = 16 * FN * GN, NT = 64 * GM * GN;
    __shared__ __align__(16) uint16_t As[TM][kLDK];
    __shared__ __align__(16) uint16_t Bs[TN][kLDK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w / GN, wn = w % GN;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kz0 = blockIdx.z * kper, kz1 = min(K, kz0 + kper);
    const int r = lane & 15, q = (lane >> 4) * 4;
    floatx4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int kc = kz0; kc < kz1; kc += kKC) {
        // stage A: TM x kKC, zero beyond M and kz1 (kz1 - kc is a multiple of 16)
        for (int v = t; v < TM * kKC / 4; v += NT) {
            int mm, kk;
            if (ak == 1) { mm = v / (kKC / 4); kk = (v % (kKC / 4)) * 4; }
            else { kk = v / (TM / 4); mm = (v % (TM / 4)) * 4; }
            uint2 x = make_uint2(0u, 0u);
            const int gm = m0 + mm, gk = kc + kk;
            if (gm < M && gk < kz1) x = *reinterpret_cast<const uint2*>(A + (size_t)gm * am + (size_t)gk * ak);
            if (ak == 1) {
                *reinterpret_cast<uint2*>(&As[mm][kk]) = x;
            } else {
                As[mm][kk] = (uint16_t)(x.x & 0xFFFFu); As[mm + 1][kk] = (uint16_t)(x.x >> 16);
                As[mm + 2][kk] = (uint16_t)(x.y & 0xFFFFu); As[mm + 3][kk] = (uint16_t)(x.y >> 16);
            }
        }
        // stage B: kKC x TN as [n][k]
        for (int v = t; v < TN * kKC / 4; v += NT) {
            int nn, kk;
            if (bk == 1) { nn = v / (kKC / 4); kk = (v % (kKC / 4)) * 4; }
            else { kk = v / (TN / 4); nn = (v % (TN / 4)) * 4; }
            uint2 x = make_uint2(0u, 0u);
            const int gn = n0 + nn, gk = kc + kk;
            if (gn < N && gk < kz1) x = *reinterpret_cast<const uint2*>(B + (size_t)gk * bk + (size_t)gn * bn);
            if (bk == 1) {
                *reinterpret_cast<uint2*>(&Bs[nn][kk]) = x;
            } else {
                Bs[nn][kk] = (uint16_t)(x.x & 0xFFFFu); Bs[nn + 1][kk] = (uint16_t)(x.x >> 16);
                Bs[nn + 2][kk] = (uint16_t)(x.y & 0xFFFFu); Bs[nn + 3][kk] = (uint16_t)(x.y >> 16);
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kKC; ks += 16) {
            bf16x4 a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i)
                a[i] = *reinterpret_cast<const bf16x4*>(&As[(wm * FM + i) * 16 + r][ks + q]);
#pragma unroll
            for (int j = 0; j < FN; ++j)
                b[j] = *reinterpret_cast<const bf16x4*>(&Bs[(wn * FN + j) * 16 + r][ks + q]);
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    float* Cz = C + (int64_t)blockIdx.z * cstride;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int gn = n0 + (wn * FN + j) * 16 + r;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int gm = m0 + (wm * FM + i) * 16 + q + e;
                if (gm < M && gn < N) Cz[(size_t)gm * ldc + gn] = acc[i][j][e];
            }
        }
}

// C[m][n] (row stride ldc) = sum over z = 0 .. nz-1 of P[z][m][n] (M x N each), in that order
__global__ void k_sum_partials(int M, int N, int nz, const float* __restrict__ P, int64_t pstride,
                               float* __restrict__ C, int ldc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * N) return;
    float s = 0.f;
    for (int z = 0; z < nz; ++z) s += P[(int64_t)z * pstride + i];
    C[(size_t)(i / N) * ldc + i % N] = s;
}

// H = tanh(Z + b) for rows < rows_valid (0 beyond), FP32 and BF16 copies
__global__ void k_bias_tanh(int M, int N, int rows_valid, const float* __restrict__ Z, const float* __restrict__ b,
                            float* __restrict__ H, uint16_t* __restrict__ Hb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * N) return;
    const int m = i / N, n = i % N;
    const float h = m < rows_valid ? tanhf(Z[i] + b[n]) : 0.0f;
    H[i] = h;
    Hb[i] = f2bf(h);
}

// Y = Z + b3 on the valid block; dY = (Y - T) / (n nout) there and 0 elsewhere; loss
__global__ void k_output(int M, int N, int rows_valid, int nout, const float* __restrict__ Z,
                         const float* __restrict__ b, const float* __restrict__ T, float* __restrict__ Y,
                         float* __restrict__ dY, uint16_t* __restrict__ dYb, double* __restrict__ loss)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < M * N) {
        const int m = i / N, n = i % N;
        const bool valid = m < rows_valid && n < nout;
        const float y = Z[i] + b[n];
        Y[i] = y;
        const float e = valid ? (y - T[i]) : 0.0f;
        const float g = e / (float)(rows_valid * nout);
        dY[i] = g;
        dYb[i] = f2bf(g);
        l = valid ? 0.5 * (double)e * (double)e / (double)(rows_valid * nout) : 0.0;
    }
    if (loss) {
        for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
        if ((threadIdx.x & 63) == 0 && l != 0.0) atomicAdd(loss, l);
    }
}

// dZ = dH * (1 - H^2), FP32 and BF16
__global__ void k_tanh_grad(int total, const float* __restrict__ dH, const float* __restrict__ H,
                            float* __restrict__ dZ, uint16_t* __restrict__ dZb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float g = dH[i] * (1.0f - H[i] * H[i]);
    dZ[i] = g;
    dZb[i] = f2bf(g);
}

// Column sums of dZ (M x N) for the bias gradients, in two ordered passes: workgroup (b, z)
// takes columns 64b .. 64b+63 and the rows of slab z (kColSlab of them), wavefront w of 16
// the slab's rows m = w (mod 16) (coalesced rows of 64 columns), and writes the slab's sum
// P[z][n] (its 16 wavefront sums added in order through LDS); k_sum_partials then adds the
// slabs in order.
constexpr int kColSlab = 128;
__global__ __launch_bounds__(1024) void k_colsum(int M, int N, const float* __restrict__ dZ, float* __restrict__ P)
{
    __shared__ float part[16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + lane;
    const int m1 = min(M, (int)(blockIdx.y + 1) * kColSlab);
    float s = 0.f;
    if (n < N)
        for (int m = blockIdx.y * kColSlab + w; m < m1; m += 16) s += dZ[(size_t)m * N + n];
    part[w][lane] = s;
    __syncthreads();
    if (w == 0 && n < N) {
        float t = 0.f;
        for (int q = 0; q < 16; ++q) t += part[q][lane];
        P[(size_t)blockIdx.y * N + n] = t;
    }
}

// Adam on the parameter array; entries with mask 0 (padding) stay zero.  Also refreshes the
// BF16 copy (same layout) the matrix products read.
__global__ void k_adam(int total, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ mo,
                       float* __restrict__ ve, const uint8_t* __restrict__ mask, float lr, float c1, float c2,
                       uint16_t* __restrict__ wb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    if (mask && !mask[i]) { w[i] = 0.f; wb[i] = 0; return; }
    const float gi = g[i];
    const float m = 0.9f * mo[i] + 0.1f * gi;
    const float v = 0.999f * ve[i] + 0.001f * gi * gi;
    mo[i] = m;
    ve[i] = v;
    const float wi = w[i] - lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
    w[i] = wi;
    wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(int total, const float* __restrict__ w, uint16_t* __restrict__ wb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) wb[i] = f2bf(w[i]);
}

inline int blocks(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }
inline int pad16(int64_t n) { return (int)((n + 15) / 16 * 16); }

}  // namespace

struct dpd_mlp {
    int nin = 0, H = 0, nout = 0, kin = 16, kout = 16;
    int device = 0;
    hipStream_t stream = nullptr;
    // FP32 parameters (padded): W1[kin][H], b1[H], W2[H][H], b2[H], W3[H][kout], b3[kout]
    float *W1 = nullptr, *b1 = nullptr, *W2 = nullptr, *b2 = nullptr, *W3 = nullptr, *b3 = nullptr;
    uint16_t* Wb = nullptr;   // BF16 copy of the parameter array (same layout)
    uint16_t *W1b = nullptr, *W2b = nullptr, *W3b = nullptr;   // its weight blocks
    // gradients and Adam moments, same shapes
    float *gW1 = nullptr, *gb1 = nullptr, *gW2 = nullptr, *gb2 = nullptr, *gW3 = nullptr, *gb3 = nullptr;
    float* mom = nullptr;     // all parameters: first moment
    float* var = nullptr;     // second moment
    uint8_t* mask = nullptr;  // 1 where a parameter is real (not padding)
    int64_t nparam_pad = 0;
    int64_t adam_t = 0;
    // activations for up to cap rows
    int cap = 0;
    uint16_t *Xb = nullptr, *H1b = nullptr, *H2b = nullptr, *dYb = nullptr, *dZ2b = nullptr, *dZ1b = nullptr;
    float *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Z3 = nullptr, *Y = nullptr, *T = nullptr;
    float *dY = nullptr, *dH2 = nullptr, *dZ2 = nullptr, *dH1 = nullptr, *dZ1 = nullptr;
    double* loss = nullptr;
    int loss_cap = 0;
    float* part = nullptr;    // split-K partial tiles
    int64_t part_cap = 0;
    std::vector<double> in_mean, in_scale, out_mean, out_scale;
    std::string err;
};

namespace {

int mfail(dpd_mlp* m, const std::string& msg)
{
    if (m) m->err = msg;
    return 3;   // DPD_ERR_HIP
}

#define MCHK(expr)                                                                     \
    do {                                                                               \
        hipError_t _st = (expr);                                                       \
        if (_st != hipSuccess) return mfail(m, std::string(#expr) + ": " + hipGetErrorString(_st)); \
    } while (0)

template <class T>
int malloc_dev(dpd_mlp* m, T** p, size_t count)
{
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    MCHK(hipMalloc((void**)p, sizeof(T) * (count ? count : 1)));
    MCHK(hipMemsetAsync(*p, 0, sizeof(T) * (count ? count : 1), m->stream));
    return 0;
}

// offsets of the padded parameter blocks inside one contiguous FP32 array
struct Layout {
    int64_t w1, b1, w2, b2, w3, b3, total;
};

Layout layout(const dpd_mlp* m)
{
    Layout L;
    L.w1 = 0;
    L.b1 = L.w1 + (int64_t)m->kin * m->H;
    L.w2 = L.b1 + m->H;
    L.b2 = L.w2 + (int64_t)m->H * m->H;
    L.w3 = L.b2 + m->H;
    L.b3 = L.w3 + (int64_t)m->H * m->kout;
    L.total = L.b3 + m->kout;
    return L;
}

int ensure_rows(dpd_mlp* m, int rows)
{
    if (rows <= m->cap) return 0;
    const int c = rows;
    const size_t h = (size_t)c * m->H, o = (size_t)c * m->kout, x = (size_t)c * m->kin;
    int rc = 0;
    if ((rc = malloc_dev(m, &m->Xb, x)) || (rc = malloc_dev(m, &m->H1b, h)) || (rc = malloc_dev(m, &m->H2b, h)) ||
        (rc = malloc_dev(m, &m->dYb, o)) || (rc = malloc_dev(m, &m->dZ2b, h)) || (rc = malloc_dev(m, &m->dZ1b, h)) ||
        (rc = malloc_dev(m, &m->Z1, h)) || (rc = malloc_dev(m, &m->H1, h)) || (rc = malloc_dev(m, &m->Z2, h)) ||
        (rc = malloc_dev(m, &m->H2, h)) || (rc = malloc_dev(m, &m->Z3, o)) || (rc = malloc_dev(m, &m->Y, o)) ||
        (rc = malloc_dev(m, &m->T, o)) || (rc = malloc_dev(m, &m->dY, o)) || (rc = malloc_dev(m, &m->dH2, h)) ||
        (rc = malloc_dev(m, &m->dZ2, h)) || (rc = malloc_dev(m, &m->dH1, h)) || (rc = malloc_dev(m, &m->dZ1, h)))
        return rc;
    m->cap = c;
    return 0;
}

// the partial-sum buffer (split-K tiles, column-sum slabs), grown to `need` floats
int ensure_part(dpd_mlp* m, int64_t need)
{
    if (need <= m->part_cap) return 0;
    int rc;
    if ((rc = malloc_dev(m, &m->part, need))) return rc;
    m->part_cap = need;
    return 0;
}

// db[n] = sum over the M rows of dZ[m][n] (N columns), in slabs of kColSlab rows
int colsum(dpd_mlp* m, int M, int N, const float* dZ, float* db)
{
    const int nslab = (M + kColSlab - 1) / kColSlab;
    int rc;
    if ((rc = ensure_part(m, (int64_t)nslab * N))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(blocks(N, 64), nslab), dim3(1024), 0, m->stream, M, N, dZ, m->part);
    hipLaunchKernelGGL(k_sum_partials, dim3(blocks(N)), dim3(256), 0, m->stream, 1, N, nslab, m->part, (int64_t)N,
                       db, N);
    return 0;
}

// tile shapes: 64 x 64 (four wavefronts of 2 x 2 fragments), 64 x 16 for N = 16 outputs,
// 16 x 64 for M = 16 (a fragment per wavefront); K split over workgroups (at least 64 each)
// when the tiles alone would leave CUs idle
int gemm(dpd_mlp* m, int M, int N, int K, const uint16_t* A, int am, int ak, const uint16_t* B, int bk, int bn,
         float* C, int ldc)
{
    if ((am != 1 && ak != 1) || (bk != 1 && bn != 1) || M % 16 || N % 16 || K % 16)
        return mfail(m, "gemm: operands must be contiguous along one dimension, sizes multiples of 16");
    int TM = 64, TN = 64;
    if (N < 64) TN = 16;
    else if (M < 64) TM = 16;
    const int gx = (N + TN - 1) / TN, gy = (M + TM - 1) / TM;
    int nz = 1;
    while (gx * gy * nz < 256 && K / (2 * nz) >= 2 * kKC) nz *= 2;
    const int kper = (K / 16 + nz - 1) / nz * 16;
    nz = (K + kper - 1) / kper;
    float* out = C;
    int ldo = ldc;
    int64_t cstride = 0;
    if (nz > 1) {
        int rc;
        if ((rc = ensure_part(m, (int64_t)nz * M * N))) return rc;
        out = m->part;
        ldo = N;
        cstride = (int64_t)M * N;
    }
    const dim3 grid(gx, gy, nz);
    if (TN == 16)
        hipLaunchKernelGGL((k_gemm_tile<1, 1, 4, 1>), grid, dim3(256), 0, m->stream, M, N, K, kper, A, am, ak, B, bk,
                           bn, out, ldo, cstride);
    else if (TM == 16)
        hipLaunchKernelGGL((k_gemm_tile<1, 1, 1, 4>), grid, dim3(256), 0, m->stream, M, N, K, kper, A, am, ak, B, bk,
                           bn, out, ldo, cstride);
    else
        hipLaunchKernelGGL((k_gemm_tile<2, 2, 2, 2>), grid, dim3(256), 0, m->stream, M, N, K, kper, A, am, ak, B, bk,
                           bn, out, ldo, cstride);
    if (nz > 1)
        hipLaunchKernelGGL(k_sum_partials, dim3(blocks((int64_t)M * N)), dim3(256), 0, m->stream, M, N, nz, m->part,
                           cstride, C, ldc);
    return 0;
}

// standardised inputs (bf16, padded to kin) and targets (fp32, padded to kout)
int upload_batch(dpd_mlp* m, int64_t n, const double* x, const double* y, int Mp)
{
    std::vector<uint16_t> xb((size_t)Mp * m->kin, 0);
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < m->nin; ++j)
            xb[(size_t)i * m->kin + j] = f2bf((float)((x[i * m->nin + j] - m->in_mean[j]) / m->in_scale[j]));
    MCHK(hipMemcpyAsync(m->Xb, xb.data(), sizeof(uint16_t) * xb.size(), hipMemcpyHostToDevice, m->stream));
    if (y) {
        std::vector<float> t((size_t)Mp * m->kout, 0.f);
        for (int64_t i = 0; i < n; ++i)
            for (int j = 0; j < m->nout; ++j)
                t[(size_t)i * m->kout + j] = (float)((y[i * m->nout + j] - m->out_mean[j]) / m->out_scale[j]);
        MCHK(hipMemcpyAsync(m->T, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice, m->stream));
    }
    return 0;
}

int forward(dpd_mlp* m, int n, int Mp, double* loss)
{
    const int H = m->H, ko = m->kout;
    int rc;
    if ((rc = gemm(m, Mp, H, m->kin, m->Xb, m->kin, 1, m->W1b, H, 1, m->Z1, H))) return rc;
    hipLaunchKernelGGL(k_bias_tanh, dim3(blocks((int64_t)Mp * H)), dim3(256), 0, m->stream, Mp, H, n, m->Z1, m->b1,
                       m->H1, m->H1b);
    if ((rc = gemm(m, Mp, H, H, m->H1b, H, 1, m->W2b, H, 1, m->Z2, H))) return rc;
    hipLaunchKernelGGL(k_bias_tanh, dim3(blocks((int64_t)Mp * H)), dim3(256), 0, m->stream, Mp, H, n, m->Z2, m->b2,
                       m->H2, m->H2b);
    if ((rc = gemm(m, Mp, ko, H, m->H2b, H, 1, m->W3b, ko, 1, m->Z3, ko))) return rc;
    hipLaunchKernelGGL(k_output, dim3(blocks((int64_t)Mp * ko)), dim3(256), 0, m->stream, Mp, ko, n, m->nout, m->Z3,
                       m->b3, m->T, m->Y, m->dY, m->dYb, loss);
    return 0;
}

int backward(dpd_mlp* m, int Mp)
{
    const int H = m->H, ko = m->kout, ki = m->kin;
    int rc;
    // dW3 = H2^T dY, db3 = colsum dY
    if ((rc = gemm(m, H, ko, Mp, m->H2b, 1, H, m->dYb, ko, 1, m->gW3, ko))) return rc;
    if ((rc = colsum(m, Mp, ko, m->dY, m->gb3))) return rc;
    // dH2 = dY W3^T
    if ((rc = gemm(m, Mp, H, ko, m->dYb, ko, 1, m->W3b, 1, ko, m->dH2, H))) return rc;
    hipLaunchKernelGGL(k_tanh_grad, dim3(blocks((int64_t)Mp * H)), dim3(256), 0, m->stream, Mp * H, m->dH2, m->H2,
                       m->dZ2, m->dZ2b);
    // dW2 = H1^T dZ2, db2
    if ((rc = gemm(m, H, H, Mp, m->H1b, 1, H, m->dZ2b, H, 1, m->gW2, H))) return rc;
    if ((rc = colsum(m, Mp, H, m->dZ2, m->gb2))) return rc;
    // dH1 = dZ2 W2^T
    if ((rc = gemm(m, Mp, H, H, m->dZ2b, H, 1, m->W2b, 1, H, m->dH1, H))) return rc;
    hipLaunchKernelGGL(k_tanh_grad, dim3(blocks((int64_t)Mp * H)), dim3(256), 0, m->stream, Mp * H, m->dH1, m->H1,
                       m->dZ1, m->dZ1b);
    // dW1 = X^T dZ1, db1
    if ((rc = gemm(m, ki, H, Mp, m->Xb, 1, ki, m->dZ1b, H, 1, m->gW1, H))) return rc;
    if ((rc = colsum(m, Mp, H, m->dZ1, m->gb1))) return rc;
    return 0;
}

void refresh_bf16(dpd_mlp* m)
{
    const Layout L = layout(m);
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(L.total)), dim3(256), 0, m->stream, (int)L.total, m->W1, m->Wb);
}

// copy between the padded device layout and the compact parameter order of the ABI
void compact_order(const dpd_mlp* m, std::vector<std::pair<int64_t, int64_t>>& map)
{
    const Layout L = layout(m);
    map.clear();
    for (int i = 0; i < m->nin; ++i)
        for (int j = 0; j < m->H; ++j) map.push_back({L.w1 + (int64_t)i * m->H + j, 0});
    for (int j = 0; j < m->H; ++j) map.push_back({L.b1 + j, 0});
    for (int64_t i = 0; i < (int64_t)m->H * m->H; ++i) map.push_back({L.w2 + i, 0});
    for (int j = 0; j < m->H; ++j) map.push_back({L.b2 + j, 0});
    for (int i = 0; i < m->H; ++i)
        for (int j = 0; j < m->nout; ++j) map.push_back({L.w3 + (int64_t)i * m->kout + j, 0});
    for (int j = 0; j < m->nout; ++j) map.push_back({L.b3 + j, 0});
    for (size_t q = 0; q < map.size(); ++q) map[q].second = (int64_t)q;
}

}  // namespace

extern "C" {

const char* dpd_mlp_last_error(const dpd_mlp* m) { return m ? m->err.c_str() : "null network"; }

int dpd_mlp_create(int32_t nin, int32_t hidden, int32_t nout, uint64_t seed, int32_t device, dpd_mlp** out)
{
    if (!out) return 1;
    *out = nullptr;
    if (nin < 1 || nin > 16 || nout < 1 || nout > 16 || hidden < 16 || hidden > 1024 || hidden % 16) return 1;
    dpd_mlp* m = new dpd_mlp();
    *out = m;
    m->nin = nin; m->H = hidden; m->nout = nout; m->device = device;
    MCHK(hipSetDevice(device));
    MCHK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    const Layout L = layout(m);
    m->nparam_pad = L.total;
    float* params = nullptr;
    int rc;
    if ((rc = malloc_dev(m, &params, L.total)) || (rc = malloc_dev(m, &m->mom, L.total)) ||
        (rc = malloc_dev(m, &m->var, L.total)) || (rc = malloc_dev(m, &m->mask, L.total)))
        return rc;
    m->W1 = params + L.w1; m->b1 = params + L.b1; m->W2 = params + L.w2;
    m->b2 = params + L.b2; m->W3 = params + L.w3; m->b3 = params + L.b3;
    float* grads = nullptr;
    if ((rc = malloc_dev(m, &grads, L.total))) return rc;
    m->gW1 = grads + L.w1; m->gb1 = grads + L.b1; m->gW2 = grads + L.w2;
    m->gb2 = grads + L.b2; m->gW3 = grads + L.w3; m->gb3 = grads + L.b3;
    if ((rc = malloc_dev(m, &m->Wb, L.total))) return rc;
    m->W1b = m->Wb + L.w1; m->W2b = m->Wb + L.w2; m->W3b = m->Wb + L.w3;
    // Glorot-uniform weights, zero biases, padding masked out
    std::mt19937_64 rng(seed);
    std::vector<float> w(L.total, 0.f);
    std::vector<uint8_t> mk(L.total, 0);
    auto fill = [&](int64_t off, int rows, int cols, int ld, int fan_in, int fan_out, bool weight) {
        const double s = std::sqrt(6.0 / (fan_in + fan_out));
        std::uniform_real_distribution<double> U(-s, s);
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j) {
                w[off + (int64_t)i * ld + j] = weight ? (float)U(rng) : 0.f;
                mk[off + (int64_t)i * ld + j] = 1;
            }
    };
    fill(L.w1, nin, hidden, hidden, nin, hidden, true);
    fill(L.b1, 1, hidden, hidden, 1, 1, false);
    fill(L.w2, hidden, hidden, hidden, hidden, hidden, true);
    fill(L.b2, 1, hidden, hidden, 1, 1, false);
    fill(L.w3, hidden, nout, m->kout, hidden, nout, true);
    fill(L.b3, 1, nout, m->kout, 1, 1, false);
    MCHK(hipMemcpyAsync(params, w.data(), sizeof(float) * L.total, hipMemcpyHostToDevice, m->stream));
    MCHK(hipMemcpyAsync(m->mask, mk.data(), L.total, hipMemcpyHostToDevice, m->stream));
    refresh_bf16(m);
    m->in_mean.assign(nin, 0.0); m->in_scale.assign(nin, 1.0);
    m->out_mean.assign(nout, 0.0); m->out_scale.assign(nout, 1.0);
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

void dpd_mlp_destroy(dpd_mlp* m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* ptrs[] = {m->W1, m->gW1, m->mom, m->var, m->mask, m->Wb, m->Xb, m->H1b, m->H2b, m->dYb,
                    m->dZ2b, m->dZ1b, m->Z1, m->H1, m->Z2, m->H2, m->Z3, m->Y, m->T, m->dY, m->dH2, m->dZ2,
                    m->dH1, m->dZ1, m->loss, m->part};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int dpd_mlp_set_scaling(dpd_mlp* m, const double* in_mean, const double* in_scale, const double* out_mean,
                        const double* out_scale)
{
    if (!m) return 1;
    if (in_mean) m->in_mean.assign(in_mean, in_mean + m->nin);
    if (in_scale) m->in_scale.assign(in_scale, in_scale + m->nin);
    if (out_mean) m->out_mean.assign(out_mean, out_mean + m->nout);
    if (out_scale) m->out_scale.assign(out_scale, out_scale + m->nout);
    for (double s : m->in_scale) if (!(s > 0)) { m->err = "scales must be positive"; return 1; }
    for (double s : m->out_scale) if (!(s > 0)) { m->err = "scales must be positive"; return 1; }
    return 0;
}

int64_t dpd_mlp_param_count(const dpd_mlp* m)
{
    if (!m) return -1;
    return (int64_t)m->nin * m->H + m->H + (int64_t)m->H * m->H + m->H + (int64_t)m->H * m->nout + m->nout;
}

int dpd_mlp_get_params(dpd_mlp* m, float* params)
{
    if (!m || !params) return 1;
    std::vector<float> w(m->nparam_pad);
    MCHK(hipMemcpyAsync(w.data(), m->W1, sizeof(float) * w.size(), hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    std::vector<std::pair<int64_t, int64_t>> map;
    compact_order(m, map);
    for (auto& p : map) params[p.second] = w[p.first];
    return 0;
}

int dpd_mlp_set_params(dpd_mlp* m, const float* params)
{
    if (!m || !params) return 1;
    std::vector<float> w(m->nparam_pad, 0.f);
    std::vector<std::pair<int64_t, int64_t>> map;
    compact_order(m, map);
    for (auto& p : map) w[p.first] = params[p.second];
    MCHK(hipMemcpyAsync(m->W1, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice, m->stream));
    refresh_bf16(m);
    MCHK(hipStreamSynchronize(m->stream));
    m->adam_t = 0;
    MCHK(hipMemsetAsync(m->mom, 0, sizeof(float) * m->nparam_pad, m->stream));
    MCHK(hipMemsetAsync(m->var, 0, sizeof(float) * m->nparam_pad, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int dpd_mlp_gradients(dpd_mlp* m, int64_t n, const double* x, const double* y, double* loss, float* grads)
{
    if (!m || n < 1 || !x || !y) return 1;
    MCHK(hipSetDevice(m->device));
    const int Mp = pad16(n);
    int rc;
    if ((rc = ensure_rows(m, Mp)) || (rc = upload_batch(m, n, x, y, Mp))) return rc;
    if (m->loss_cap < 1) {
        if ((rc = malloc_dev(m, &m->loss, 1))) return rc;
        m->loss_cap = 1;
    }
    MCHK(hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
    if ((rc = forward(m, (int)n, Mp, m->loss)) || (rc = backward(m, Mp))) return rc;
    MCHK(hipGetLastError());
    double l = 0;
    MCHK(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    std::vector<float> g(m->nparam_pad);
    MCHK(hipMemcpyAsync(g.data(), m->gW1, sizeof(float) * g.size(), hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    if (loss) *loss = l;
    if (grads) {
        std::vector<std::pair<int64_t, int64_t>> map;
        compact_order(m, map);
        for (auto& p : map) grads[p.second] = g[p.first];
    }
    return 0;
}

int dpd_mlp_train(dpd_mlp* m, int64_t n, const double* x, const double* y, int32_t iters, double lr, double* loss_out)
{
    if (!m || n < 1 || !x || !y || iters < 0) return 1;
    MCHK(hipSetDevice(m->device));
    const int Mp = pad16(n);
    int rc;
    if ((rc = ensure_rows(m, Mp)) || (rc = upload_batch(m, n, x, y, Mp))) return rc;
    if (m->loss_cap < iters) {
        if ((rc = malloc_dev(m, &m->loss, iters))) return rc;
        m->loss_cap = iters;
    }
    MCHK(hipMemsetAsync(m->loss, 0, sizeof(double) * std::max(iters, 1), m->stream));
    const int64_t tot = m->nparam_pad;
    for (int32_t it = 0; it < iters; ++it) {
        if ((rc = forward(m, (int)n, Mp, m->loss + it)) || (rc = backward(m, Mp))) return rc;
        m->adam_t += 1;
        const float c1 = 1.0f - std::pow(0.9f, (float)m->adam_t);
        const float c2 = 1.0f - std::pow(0.999f, (float)m->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(blocks(tot)), dim3(256), 0, m->stream, (int)tot, m->W1, m->gW1, m->mom,
                           m->var, m->mask, (float)lr, c1, c2, m->Wb);
    }
    MCHK(hipGetLastError());
    if (loss_out && iters > 0)
        MCHK(hipMemcpyAsync(loss_out, m->loss, sizeof(double) * iters, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int dpd_mlp_predict(dpd_mlp* m, int64_t n, const double* x, double* y)
{
    if (!m || n < 1 || !x || !y) return 1;
    MCHK(hipSetDevice(m->device));
    const int Mp = pad16(n);
    int rc;
    if ((rc = ensure_rows(m, Mp)) || (rc = upload_batch(m, n, x, nullptr, Mp))) return rc;
    if ((rc = forward(m, (int)n, Mp, nullptr))) return rc;
    MCHK(hipGetLastError());
    std::vector<float> out((size_t)Mp * m->kout);
    MCHK(hipMemcpyAsync(out.data(), m->Y, sizeof(float) * out.size(), hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < m->nout; ++j)
            y[i * m->nout + j] = m->out_mean[j] + m->out_scale[j] * (double)out[(size_t)i * m->kout + j];
    return 0;
}

}  // extern "C"
