// k_field.hip -- Poisson field solve as dense transform-matrix products on MFMA.
//
// Reference: the separable DCT field of vck5000/host/src/AIEplace.cpp:752-934
// (compute_a_uv_DCT / compute_eField_DCT with DCT_naive/IDCT_naive/IDXST_naive,
// DCT.cpp:24-82) and the AIE DCT/IDCT/IDXST graphs (DCTGraph.h, IDCTGraph.h,
// IDXSTGraph.h), in the RePlAce scaling of python/AIEmath/fft_replace/fft.cpp:118-212.
// Each 1-D transform pass over all rows/columns is one GEMM with a cos/sin
// matrix; the spectral scaling is fused into the epilogue of the second GEMM.
//
// Every product is C[m][n] = sum_k A[m][k] * Bt[n][k] ("NT": K contiguous in
// both operands), A a constant transform matrix and Bt the data:
//   G1  T1t[v][i] = sum_j Cy[v][j]  rho[i][j]           (Cy[v][j] = cos(pi v (j+1/2)/My))
//   G2  A[u][v]   = sum_i Cx[u][i]  T1t[v][i]  -> spectral -> Ehx[u][v], Ehy[u][v]
//   G3  R1t[j][u] = sum_v CyT[j][v] Ehx[u][v];  R2t[j][u] = sum_v SyT[j][v] Ehy[u][v]
//   G4  Ex[i][j]  = sum_u SxT[i][u] R1t[j][u];  Ey[i][j]  = sum_u CxT[i][u] R2t[j][u]
// Map layout: rho[i*My + j], i = x bin, j = y bin.
//
// Precision forms:
//   FP32     v_mfma_f32_32x32x2_f32
//   BF16X3   each operand x = hi + lo in BF16; hi*hi + hi*lo + lo*hi on
//            v_mfma_f32_32x32x16_bf16 with FP32 accumulation
//   FP16X3   the same split in FP16 on v_mfma_f32_32x32x16_f16; the data operand
//            is scaled by a power of two from its running max |value| so hi and
//            lo stay in range (each producer tracks the max of its output)
//   FP64     reference solve (FMA, FP64 throughout) used to report the error of
//            the other forms
// The constant matrices are split once at set-up; the data operand is split
// while it is staged into LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "ep_batch.h"
#include "ep_ctx.h"

namespace ep {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int BM = 64, BN = 64;
constexpr float kF16Top = 14.0f;  // scaled data max lies in [2^14, 2^15)

struct GemmArgs {
    const float *A[2];          // FP32 constant operand [M][K]
    const uint16_t *Ah[2];      // split constant operand, hi plane (BF16 or FP16 bits)
    const uint16_t *Al[2];      //                          lo plane
    const float *B[2];          // data operand [N][K], FP32
    float *C[2];                // [M][N]
    float *C2[2];               // second output of the spectral epilogue
    int M, N, K;
    float scale;  // spectral: 4/(Mx My) / bin area
    float wxs;    // wx = wxs * u
    float wys;    // wy = wys * v
    const uint32_t *bmax;       // FP16X3: max |data operand| (float bits)
    uint32_t *cmax;             // FP16X3: running max |output| of this product
    const float *bbias;         // nullable: constant subtracted from the data operand
    const int32_t *skip;        // batched: nonzero when the instance has finished
};

__device__ inline float spectral_val(const GemmArgs &a, int row, int col, float v, float *o2)
{
    float val = v * a.scale;
    if (row == 0) val *= 0.5f;
    if (col == 0) val *= 0.5f;
    const float wx = a.wxs * (float)row, wy = a.wys * (float)col;
    const float p = (row == 0 && col == 0) ? 0.0f : val / (wx * wx + wy * wy);
    *o2 = p * wy;
    return p * wx;
}

// Stores one output element; returns max |written value| for max tracking.
__device__ inline float store_out(const GemmArgs &a, int z, int epi, int row, int col, float v)
{
    const size_t o = (size_t)row * a.N + col;
    if (epi == 0) {
        a.C[z][o] = v;
        return fabsf(v);
    }
    float v2;
    const float v1 = spectral_val(a, row, col, v, &v2);
    a.C[z][o] = v1;
    a.C2[z][o] = v2;
    return fmaxf(fabsf(v1), fabsf(v2));
}

// Block max of m, one atomic per block (positive floats order like their bits).
__device__ inline void track_max(uint32_t *slot, float m)
{
    __shared__ float wmax[4];
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float b = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        if (b > 0.0f) atomicMax(slot, __float_as_uint(b));
    }
}

// ---------------------------------------------------------------- FP32

constexpr int BK32 = 16;

// 64x64 block tile, 4 waves of 32x32; global tiles [64][16] of A and Bt are
// staged transposed into LDS ([k][m], [k][n]) for the 32x32x2 operand layout.
template <int EPI>
__device__ inline void gemm_f32_tile(const GemmArgs &a, int z, int m0, int n0)
{
    const float *__restrict__ A = a.A[z];
    const float *__restrict__ B = a.B[z];
    const int M = a.M, N = a.N, K = a.K;
    __shared__ float As[BK32][BM + 4];
    __shared__ float Bs[BK32][BN + 4];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    const int lr = t >> 2, lk = (t & 3) * 4;
    const bool a_ok = (m0 + lr) < M, b_ok = (n0 + lr) < N;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

    const float bias = (a.bbias && b_ok) ? *a.bbias : 0.0f;
    float4 av = make_float4(0, 0, 0, 0), bv = make_float4(0, 0, 0, 0);
    if (a_ok) av = *reinterpret_cast<const float4 *>(&A[(size_t)(m0 + lr) * K + lk]);
    if (b_ok) bv = *reinterpret_cast<const float4 *>(&B[(size_t)(n0 + lr) * K + lk]);
    for (int k0 = 0; k0 < K; k0 += BK32) {
        __syncthreads();
        As[lk + 0][lr] = av.x; As[lk + 1][lr] = av.y; As[lk + 2][lr] = av.z; As[lk + 3][lr] = av.w;
        Bs[lk + 0][lr] = bv.x - bias; Bs[lk + 1][lr] = bv.y - bias;
        Bs[lk + 2][lr] = bv.z - bias; Bs[lk + 3][lr] = bv.w - bias;
        __syncthreads();
        const int kn = k0 + BK32;
        if (kn < K) {  // prefetch the next tiles while the MFMAs run
            if (a_ok) av = *reinterpret_cast<const float4 *>(&A[(size_t)(m0 + lr) * K + kn + lk]);
            if (b_ok) bv = *reinterpret_cast<const float4 *>(&B[(size_t)(n0 + lr) * K + kn + lk]);
        }
#pragma unroll
        for (int kk = 0; kk < BK32; kk += 2) {
            const float x = As[kk + (lane >> 5)][wm + (lane & 31)];
            const float y = Bs[kk + (lane >> 5)][wn + (lane & 31)];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x, y, acc, 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const int col = n0 + wn + (lane & 31);
        if (row < M && col < N) store_out(a, z, EPI, row, col, acc[i]);
    }
}

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f32(GemmArgs a)
{
    gemm_f32_tile<EPI>(a, blockIdx.z, blockIdx.y * BM, blockIdx.x * BN);
}

// ---------------------------------------------------------------- split 16-bit

constexpr int BK16 = 32;
constexpr int LDK = BK16 + 8;  // padded LDS row (16-bit elements)

__device__ inline uint16_t bf16_rne(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <bool F16>
__device__ inline void split16(float x, uint16_t &h, uint16_t &l)
{
    if (F16) {
        const _Float16 hh = (_Float16)x;
        const _Float16 ll = (_Float16)(x - (float)hh);
        h = __builtin_bit_cast(uint16_t, hh);
        l = __builtin_bit_cast(uint16_t, ll);
    } else {
        h = bf16_rne(x);
        l = bf16_rne(x - bf16_to_f(h));
    }
}

template <bool F16>
__device__ inline f32x16 mfma16(const uint16_t *pa, const uint16_t *pb, f32x16 acc)
{
    if (F16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(*reinterpret_cast<const f16x8 *>(pa),
                                                     *reinterpret_cast<const f16x8 *>(pb), acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(*reinterpret_cast<const bf16x8 *>(pa),
                                                  *reinterpret_cast<const bf16x8 *>(pb), acc, 0, 0, 0);
}

// 64x64 block tile, 4 waves of 32x32, K step 32 (two 32x32x16 MFMA sub-steps,
// three products each).  LDS rows are [m or n][k] with k contiguous, so each
// lane reads its 8 consecutive k values with one 16-byte load.
template <int EPI, bool F16>
__device__ inline void gemm_split_tile(const GemmArgs &a, int z, int m0, int n0)
{
    const uint16_t *__restrict__ Ah = a.Ah[z];
    const uint16_t *__restrict__ Al = a.Al[z];
    const float *__restrict__ B = a.B[z];
    const int M = a.M, N = a.N, K = a.K;
    __shared__ __attribute__((aligned(16))) uint16_t sAh[BM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sAl[BM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBh[BN][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBl[BN][LDK];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    const int lr = t >> 2, lk = (t & 3) * 8;  // row and first k of this thread's 8-element chunk
    const bool a_ok = (m0 + lr) < M, b_ok = (n0 + lr) < N;
    // FP16: power-of-two scale of the data operand (exact; undone in the epilogue)
    float bscale = 1.0f;
    if (F16) {
        const float bm = __uint_as_float(*a.bmax);
        if (bm > 0.0f && isfinite(bm)) bscale = ldexpf(1.0f, (int)kF16Top - ilogbf(bm));
    }
    const float bias = (a.bbias && b_ok) ? *a.bbias : 0.0f;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

    uint4 ahv = make_uint4(0, 0, 0, 0), alv = make_uint4(0, 0, 0, 0);
    float4 b0 = make_float4(0, 0, 0, 0), b1 = make_float4(0, 0, 0, 0);
    auto load = [&](int k0) {
        if (a_ok) {
            ahv = *reinterpret_cast<const uint4 *>(&Ah[(size_t)(m0 + lr) * K + k0 + lk]);
            alv = *reinterpret_cast<const uint4 *>(&Al[(size_t)(m0 + lr) * K + k0 + lk]);
        }
        if (b_ok) {
            const float *src = &B[(size_t)(n0 + lr) * K + k0 + lk];
            b0 = *reinterpret_cast<const float4 *>(src);
            b1 = *reinterpret_cast<const float4 *>(src + 4);
        }
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += BK16) {
        __syncthreads();
        *reinterpret_cast<uint4 *>(&sAh[lr][lk]) = ahv;
        *reinterpret_cast<uint4 *>(&sAl[lr][lk]) = alv;
        {
            const float v[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
            uint16_t h[8], l[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) split16<F16>(F16 ? (v[e] - bias) * bscale : v[e] - bias, h[e], l[e]);
            uint4 hv, lv;
            hv.x = h[0] | ((uint32_t)h[1] << 16); hv.y = h[2] | ((uint32_t)h[3] << 16);
            hv.z = h[4] | ((uint32_t)h[5] << 16); hv.w = h[6] | ((uint32_t)h[7] << 16);
            lv.x = l[0] | ((uint32_t)l[1] << 16); lv.y = l[2] | ((uint32_t)l[3] << 16);
            lv.z = l[4] | ((uint32_t)l[5] << 16); lv.w = l[6] | ((uint32_t)l[7] << 16);
            *reinterpret_cast<uint4 *>(&sBh[lr][lk]) = hv;
            *reinterpret_cast<uint4 *>(&sBl[lr][lk]) = lv;
        }
        __syncthreads();
        if (k0 + BK16 < K) load(k0 + BK16);
#pragma unroll
        for (int ks = 0; ks < BK16; ks += 16) {
            const int kk = ks + 8 * (lane >> 5);
            const uint16_t *xh = &sAh[wm + (lane & 31)][kk], *xl = &sAl[wm + (lane & 31)][kk];
            const uint16_t *yh = &sBh[wn + (lane & 31)][kk], *yl = &sBl[wn + (lane & 31)][kk];
            acc = mfma16<F16>(xh, yh, acc);
            acc = mfma16<F16>(xh, yl, acc);
            acc = mfma16<F16>(xl, yh, acc);
        }
    }
    const float inv = 1.0f / bscale;
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const int col = n0 + wn + (lane & 31);
        if (row < M && col < N) m = fmaxf(m, store_out(a, z, EPI, row, col, acc[i] * inv));
    }
    if (F16 && a.cmax) track_max(a.cmax, m);
}

template <int EPI, bool F16>
__global__ __launch_bounds__(256) void k_gemm_split(GemmArgs a)
{
    gemm_split_tile<EPI, F16>(a, blockIdx.z, blockIdx.y * BM, blockIdx.x * BN);
}

// Batched products: block b runs tile tiles[b] = (instance, m0, n0, z) with the
// instance's argument block; finished instances are skipped (uniform per block).
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f32_b(const GemmArgs *__restrict__ args,
                                                    const int4 *__restrict__ tiles)
{
    const int4 t = tiles[blockIdx.x];
    const GemmArgs &a = args[t.x];
    if (a.skip && *a.skip) return;
    gemm_f32_tile<EPI>(a, t.w, t.y, t.z);
}

template <int EPI, bool F16>
__global__ __launch_bounds__(256) void k_gemm_split_b(const GemmArgs *__restrict__ args,
                                                      const int4 *__restrict__ tiles)
{
    const int4 t = tiles[blockIdx.x];
    const GemmArgs &a = args[t.x];
    if (a.skip && *a.skip) return;
    gemm_split_tile<EPI, F16>(a, t.w, t.y, t.z);
}

// ---------------------------------------------------------------- FP64 reference

struct GemmArgs64 {
    const double *A[2];
    const float *Bf;     // data operand as FP32 (first product), else Bd
    const float *bbias;  // with Bf: constant subtracted from it
    const double *Bd[2];
    double *C[2];
    double *C2[2];
    int M, N, K;
    double scale, wxs, wys;
};

// 64x64 tile, 256 threads, each 4x4 outputs; plain FP64 FMA.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f64(GemmArgs64 a)
{
    const int z = blockIdx.z;
    const int M = a.M, N = a.N, K = a.K;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    __shared__ double As[16][65];
    __shared__ double Bs[16][65];
    const int t = threadIdx.x, tm = (t >> 4) * 4, tn = (t & 15) * 4;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int e = t; e < 64 * 16; e += 256) {
            const int r = e / 16, k = e % 16;
            As[k][r] = (m0 + r < M) ? a.A[z][(size_t)(m0 + r) * K + k0 + k] : 0.0;
            double bv = 0.0;
            if (n0 + r < N)
                bv = a.Bf ? (double)a.Bf[(size_t)(n0 + r) * K + k0 + k] - (double)*a.bbias
                          : a.Bd[z][(size_t)(n0 + r) * K + k0 + k];
            Bs[k][r] = bv;
        }
        __syncthreads();
        for (int k = 0; k < 16; ++k)
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(As[k][tm + i], Bs[k][tn + j], acc[i][j]);
        __syncthreads();
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const int row = m0 + tm + i, col = n0 + tn + j;
            if (row >= M || col >= N) continue;
            const size_t o = (size_t)row * N + col;
            if (EPI == 0) {
                a.C[z][o] = acc[i][j];
            } else {
                double val = acc[i][j] * a.scale;
                if (row == 0) val *= 0.5;
                if (col == 0) val *= 0.5;
                const double wx = a.wxs * row, wy = a.wys * col;
                const double p = (row == 0 && col == 0) ? 0.0 : val / (wx * wx + wy * wy);
                a.C[z][o] = p * wx;
                a.C2[z][o] = p * wy;
            }
        }
}

__global__ void k_clear_max(uint32_t *m, int n)
{
    if (threadIdx.x < n) m[threadIdx.x] = 0u;
}

// Block partials of max|ex - exd|, max|exd|, max|ey - eyd|, max|eyd|.
__global__ __launch_bounds__(256) void k_field_err(const float *__restrict__ ex,
                                                   const float *__restrict__ ey,
                                                   const double *__restrict__ exd,
                                                   const double *__restrict__ eyd, int nb,
                                                   double *__restrict__ part)
{
    double v[4] = {0, 0, 0, 0};
    for (int b = blockIdx.x * 256 + threadIdx.x; b < nb; b += gridDim.x * 256) {
        v[0] = fmax(v[0], fabs((double)ex[b] - exd[b]));
        v[1] = fmax(v[1], fabs(exd[b]));
        v[2] = fmax(v[2], fabs((double)ey[b] - eyd[b]));
        v[3] = fmax(v[3], fabs(eyd[b]));
    }
    __shared__ double red[4][4];
    for (int k = 0; k < 4; ++k) {
        for (int d = 32; d >= 1; d >>= 1) v[k] = fmax(v[k], __shfl_xor(v[k], d, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double m = red[0][threadIdx.x];
        for (int w = 1; w < 4; ++w) m = fmax(m, red[w][threadIdx.x]);
        part[blockIdx.x * 4 + threadIdx.x] = m;
    }
}

template <int EPI>
void launch_split(const GemmArgs &a, dim3 grid, int precision, hipStream_t s)
{
    if (precision == EP_PREC_FP16X3)
        hipLaunchKernelGGL((k_gemm_split<EPI, true>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_gemm_split<EPI, false>), grid, dim3(256), 0, s, a);
}

void gemm(const GemmArgs &a, int batch, int epi, int precision, hipStream_t s)
{
    dim3 grid((a.N + BN - 1) / BN, (a.M + BM - 1) / BM, batch);
    if (precision == EP_PREC_BF16X3 || precision == EP_PREC_FP16X3) {
        if (epi == 0)
            launch_split<0>(a, grid, precision, s);
        else
            launch_split<1>(a, grid, precision, s);
    } else {
        if (epi == 0)
            hipLaunchKernelGGL(k_gemm_f32<0>, grid, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(k_gemm_f32<1>, grid, dim3(256), 0, s, a);
    }
}

// T[r][c] = cos|sin(pi * f * (x + 1/2) / M) with (f, x) = (r, c) or, transposed, (c, r).
std::vector<double> trig(int M, bool sine, bool transpose)
{
    std::vector<double> t((size_t)M * M);
    for (int f = 0; f < M; ++f)
        for (int x = 0; x < M; ++x) {
            const double ang = M_PI * (double)f * ((double)x + 0.5) / (double)M;
            t[transpose ? (size_t)x * M + f : (size_t)f * M + x] = sine ? std::sin(ang) : std::cos(ang);
        }
    return t;
}

uint16_t bf16_rne_host(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float bf16_to_f_host(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

uint16_t f16_host(float x)
{
    const _Float16 h = (_Float16)x;
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

float f16_to_f_host(uint16_t u)
{
    _Float16 h;
    std::memcpy(&h, &u, 2);
    return (float)h;
}

}  // namespace

// Matrix slots (all copies share the index).
enum { MAT_CY = 0, MAT_CX, MAT_CYT, MAT_SYT, MAT_SXT, MAT_CXT, MAT_N };

struct FieldMats {
    float *f32[MAT_N];
    uint16_t *hi[MAT_N], *lo[MAT_N];       // BF16 split
    uint16_t *hi16[MAT_N], *lo16[MAT_N];   // FP16 split
    double *f64[MAT_N];                    // FP64 reference (built on first use)
    uint32_t *fmax;                        // FP16X3 running maxima: rho, T1t, Eh, Rt
    double *tA64, *tB64, *exd, *eyd;       // FP64 reference scratch
    double *err_part;                      // k_field_err partials
};

constexpr int kErrBlocks = 256;

static bool matrix_desc(const Ctx &c, int m, int &M, bool &sine, bool &tr)
{
    const int Ms[MAT_N] = {c.My, c.Mx, c.My, c.My, c.Mx, c.Mx};
    const bool si[MAT_N] = {false, false, false, true, true, false};
    const bool t[MAT_N] = {false, false, true, true, true, true};
    M = Ms[m];
    sine = si[m];
    tr = t[m];
    return true;
}

ep_status field_init(Ctx &c)
{
    FieldMats *fm = new (std::nothrow) FieldMats();
    if (!fm) return fail(EP_ENOMEM, "field_init: host allocation failed");
    c.split_mats = fm;
    for (int m = 0; m < MAT_N; ++m) {
        int M;
        bool sine, tr;
        matrix_desc(c, m, M, sine, tr);
        const std::vector<double> d = trig(M, sine, tr);
        std::vector<float> f(d.size());
        std::vector<uint16_t> h(d.size()), l(d.size()), h16(d.size()), l16(d.size());
        for (size_t k = 0; k < d.size(); ++k) {
            f[k] = (float)d[k];
            h[k] = bf16_rne_host(f[k]);
            l[k] = bf16_rne_host(f[k] - bf16_to_f_host(h[k]));
            h16[k] = f16_host(f[k]);
            l16[k] = f16_host(f[k] - f16_to_f_host(h16[k]));
        }
        if (hipMalloc(&fm->f32[m], f.size() * 4) != hipSuccess ||
            hipMalloc(&fm->hi[m], h.size() * 2) != hipSuccess ||
            hipMalloc(&fm->lo[m], l.size() * 2) != hipSuccess ||
            hipMalloc(&fm->hi16[m], h.size() * 2) != hipSuccess ||
            hipMalloc(&fm->lo16[m], l.size() * 2) != hipSuccess)
            return fail(EP_ENOMEM, "field_init: hipMalloc failed");
        if (hipMemcpy(fm->f32[m], f.data(), f.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(fm->hi[m], h.data(), h.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(fm->lo[m], l.data(), l.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(fm->hi16[m], h16.data(), h16.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(fm->lo16[m], l16.data(), l16.size() * 2, hipMemcpyHostToDevice) != hipSuccess)
            return fail(EP_EDEVICE, "field_init: upload failed");
    }
    // [0,4): split-FP16 operand maxima; [4]: field map mean (float)
    if (hipMalloc(&fm->fmax, 8 * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(fm->fmax, 0, 8 * sizeof(uint32_t)) != hipSuccess)
        return fail(EP_ENOMEM, "field_init: hipMalloc failed");
    c.field_max = fm->fmax;
    c.map_mean = reinterpret_cast<float *>(fm->fmax + 4);
    return EP_OK;
}

void field_free(Ctx &c)
{
    FieldMats *fm = static_cast<FieldMats *>(c.split_mats);
    if (!fm) return;
    for (int m = 0; m < MAT_N; ++m) {
        void *p[] = {fm->f32[m], fm->hi[m], fm->lo[m], fm->hi16[m], fm->lo16[m], fm->f64[m]};
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
    void *p[] = {fm->fmax, fm->tA64, fm->tB64, fm->exd, fm->eyd, fm->err_part};
    for (void *q : p)
        if (q) (void)hipFree(q);
    delete fm;
    c.split_mats = nullptr;
    c.field_max = nullptr;
    c.map_mean = nullptr;
}

void launch_field(Ctx &c, int precision, hipStream_t s)
{
    FieldMats *fm = static_cast<FieldMats *>(c.split_mats);
    const size_t nb = (size_t)c.Mx * c.My;
    const bool f16 = precision == EP_PREC_FP16X3;
    GemmArgs a{};
    a.scale = (float)(4.0 / ((double)c.Mx * c.My) / ((double)c.bw * c.bh));
    a.wxs = (float)(M_PI / c.Mx);
    a.wys = (float)(M_PI / c.My * ((double)c.bw / c.bh));
    auto setA = [&](int z, int m) {
        a.A[z] = fm->f32[m];
        a.Ah[z] = f16 ? fm->hi16[m] : fm->hi[m];
        a.Al[z] = f16 ? fm->lo16[m] : fm->lo[m];
    };
    if (f16) hipLaunchKernelGGL(k_clear_max, dim3(1), dim3(64), 0, s, fm->fmax + 1, 3);
    // G1: T1t[v][i] = Cy * rho^T
    a.M = c.My; a.N = c.Mx; a.K = c.My;
    setA(0, MAT_CY); a.B[0] = c.map_field; a.C[0] = c.tA;
    a.bmax = fm->fmax + 0; a.cmax = f16 ? fm->fmax + 1 : nullptr;
    a.bbias = c.map_mean;  // the DC mode carries no field; removing it keeps FP32 sums small
    gemm(a, 1, 0, precision, s);
    a.bbias = nullptr;
    // G2: A[u][v] = Cx * T1t^T, spectral -> Ehx (tB), Ehy (tB + nb)
    a.M = c.Mx; a.N = c.My; a.K = c.Mx;
    setA(0, MAT_CX); a.B[0] = c.tA; a.C[0] = c.tB; a.C2[0] = c.tB + nb;
    a.bmax = fm->fmax + 1; a.cmax = f16 ? fm->fmax + 2 : nullptr;
    gemm(a, 1, 1, precision, s);
    // G3: R1t[j][u] = CyT * Ehx^T, R2t = SyT * Ehy^T
    a.M = c.My; a.N = c.Mx; a.K = c.My;
    setA(0, MAT_CYT); a.B[0] = c.tB; a.C[0] = c.tA;
    setA(1, MAT_SYT); a.B[1] = c.tB + nb; a.C[1] = c.tA + nb;
    a.bmax = fm->fmax + 2; a.cmax = f16 ? fm->fmax + 3 : nullptr;
    gemm(a, 2, 0, precision, s);
    // G4: Ex[i][j] = SxT * R1t^T, Ey = CxT * R2t^T
    a.M = c.Mx; a.N = c.My; a.K = c.Mx;
    setA(0, MAT_SXT); a.B[0] = c.tA; a.C[0] = c.ex;
    setA(1, MAT_CXT); a.B[1] = c.tA + nb; a.C[1] = c.ey;
    a.bmax = fm->fmax + 3; a.cmax = nullptr;
    gemm(a, 2, 0, precision, s);
}

// FP64 field of the current map into the reference scratch, then the max-norm
// relative error of ex/ey against it.
ep_status field_error(Ctx &c, double *err_ex, double *err_ey)
{
    FieldMats *fm = static_cast<FieldMats *>(c.split_mats);
    const size_t nb = (size_t)c.Mx * c.My;
    if (!fm->f64[0]) {
        for (int m = 0; m < MAT_N; ++m) {
            int M;
            bool sine, tr;
            matrix_desc(c, m, M, sine, tr);
            const std::vector<double> d = trig(M, sine, tr);
            if (hipMalloc(&fm->f64[m], d.size() * 8) != hipSuccess ||
                hipMemcpy(fm->f64[m], d.data(), d.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
                return fail(EP_ENOMEM, "field_error: FP64 matrices");
        }
        if (hipMalloc(&fm->tA64, 2 * nb * 8) != hipSuccess || hipMalloc(&fm->tB64, 2 * nb * 8) != hipSuccess ||
            hipMalloc(&fm->exd, nb * 8) != hipSuccess || hipMalloc(&fm->eyd, nb * 8) != hipSuccess ||
            hipMalloc(&fm->err_part, kErrBlocks * 4 * 8) != hipSuccess)
            return fail(EP_ENOMEM, "field_error: FP64 scratch");
    }
    hipStream_t s = c.stream;
    GemmArgs64 a{};
    a.scale = 4.0 / ((double)c.Mx * c.My) / ((double)c.bw * c.bh);
    a.wxs = M_PI / c.Mx;
    a.wys = M_PI / c.My * ((double)c.bw / c.bh);
    auto run = [&](int batch, int epi) {
        dim3 grid((a.N + 63) / 64, (a.M + 63) / 64, batch);
        if (epi == 0)
            hipLaunchKernelGGL(k_gemm_f64<0>, grid, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(k_gemm_f64<1>, grid, dim3(256), 0, s, a);
    };
    a.M = c.My; a.N = c.Mx; a.K = c.My;
    a.A[0] = fm->f64[MAT_CY]; a.Bf = c.map_field; a.bbias = c.map_mean; a.C[0] = fm->tA64;
    run(1, 0);
    a.Bf = nullptr;
    a.M = c.Mx; a.N = c.My; a.K = c.Mx;
    a.A[0] = fm->f64[MAT_CX]; a.Bd[0] = fm->tA64; a.C[0] = fm->tB64; a.C2[0] = fm->tB64 + nb;
    run(1, 1);
    a.M = c.My; a.N = c.Mx; a.K = c.My;
    a.A[0] = fm->f64[MAT_CYT]; a.Bd[0] = fm->tB64; a.C[0] = fm->tA64;
    a.A[1] = fm->f64[MAT_SYT]; a.Bd[1] = fm->tB64 + nb; a.C[1] = fm->tA64 + nb;
    run(2, 0);
    a.M = c.Mx; a.N = c.My; a.K = c.Mx;
    a.A[0] = fm->f64[MAT_SXT]; a.Bd[0] = fm->tA64; a.C[0] = fm->exd;
    a.A[1] = fm->f64[MAT_CXT]; a.Bd[1] = fm->tA64 + nb; a.C[1] = fm->eyd;
    run(2, 0);
    hipLaunchKernelGGL(k_field_err, dim3(kErrBlocks), dim3(256), 0, s, c.ex, c.ey, fm->exd, fm->eyd,
                       (int)nb, fm->err_part);
    if (hipGetLastError() != hipSuccess) return fail(EP_EDEVICE, "field_error: launch failed");
    std::vector<double> part(kErrBlocks * 4);
    if (hipMemcpyAsync(part.data(), fm->err_part, part.size() * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(EP_EDEVICE, "field_error: copy failed");
    double m[4] = {0, 0, 0, 0};
    for (int b = 0; b < kErrBlocks; ++b)
        for (int k = 0; k < 4; ++k) m[k] = std::max(m[k], part[b * 4 + k]);
    if (err_ex) *err_ex = m[1] > 0 ? m[0] / m[1] : 0.0;
    if (err_ey) *err_ey = m[3] > 0 ? m[2] / m[3] : 0.0;
    return EP_OK;
}

// ---------------------------------------------------------------- batched solve

namespace {

__global__ void k_clear_max_b(uint32_t *fmax, int n_inst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_inst) {
        fmax[4 * i + 1] = 0u;
        fmax[4 * i + 2] = 0u;
        fmax[4 * i + 3] = 0u;
    }
}

struct MatSet {  // the six transform matrices of one grid size
    float *f32[MAT_N] = {};
    uint16_t *hi[MAT_N] = {}, *lo[MAT_N] = {}, *hi16[MAT_N] = {}, *lo16[MAT_N] = {};
};

}  // namespace

struct BatchField {
    std::vector<int> sizes;                  // distinct M
    std::vector<MatSet> sets;
    GemmArgs *d_args[3][4] = {};             // [precision][stage] per-instance argument blocks
    int4 *d_tiles[4] = {};
    int n_tiles[4] = {};
    uint32_t *fmax = nullptr;                // [4 * n_inst]
    int n_inst = 0;
};

ep_status batch_field_init(BatchField *&out, const std::vector<BatchFieldInst> &inst)
{
    BatchField *bf = new (std::nothrow) BatchField();
    if (!bf) return fail(EP_ENOMEM, "batch field: host allocation failed");
    out = bf;
    bf->n_inst = (int)inst.size();
    for (const BatchFieldInst &d : inst)
        if (std::find(bf->sizes.begin(), bf->sizes.end(), d.M) == bf->sizes.end()) bf->sizes.push_back(d.M);
    bf->sets.resize(bf->sizes.size());
    for (size_t si = 0; si < bf->sizes.size(); ++si) {
        const int M = bf->sizes[si];
        for (int m = 0; m < MAT_N; ++m) {
            const bool sine = (m == MAT_SYT || m == MAT_SXT);
            const bool tr = !(m == MAT_CY || m == MAT_CX);
            const std::vector<double> d = trig(M, sine, tr);
            std::vector<float> f(d.size());
            std::vector<uint16_t> h(d.size()), l(d.size()), h16(d.size()), l16(d.size());
            for (size_t k = 0; k < d.size(); ++k) {
                f[k] = (float)d[k];
                h[k] = bf16_rne_host(f[k]);
                l[k] = bf16_rne_host(f[k] - bf16_to_f_host(h[k]));
                h16[k] = f16_host(f[k]);
                l16[k] = f16_host(f[k] - f16_to_f_host(h16[k]));
            }
            MatSet &s = bf->sets[si];
            if (hipMalloc(&s.f32[m], f.size() * 4) != hipSuccess || hipMalloc(&s.hi[m], h.size() * 2) != hipSuccess ||
                hipMalloc(&s.lo[m], l.size() * 2) != hipSuccess || hipMalloc(&s.hi16[m], h.size() * 2) != hipSuccess ||
                hipMalloc(&s.lo16[m], l.size() * 2) != hipSuccess)
                return fail(EP_ENOMEM, "batch field: hipMalloc failed");
            if (hipMemcpy(s.f32[m], f.data(), f.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(s.hi[m], h.data(), h.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(s.lo[m], l.data(), l.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(s.hi16[m], h16.data(), h16.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(s.lo16[m], l16.data(), l16.size() * 2, hipMemcpyHostToDevice) != hipSuccess)
                return fail(EP_EDEVICE, "batch field: upload failed");
        }
    }
    const int n = bf->n_inst;
    if (hipMalloc(&bf->fmax, sizeof(uint32_t) * 4 * std::max(1, n)) != hipSuccess ||
        hipMemset(bf->fmax, 0, sizeof(uint32_t) * 4 * std::max(1, n)) != hipSuccess)
        return fail(EP_ENOMEM, "batch field: hipMalloc failed");
    // argument blocks per precision and stage, and the tile lists
    const int precs[3] = {EP_PREC_FP32, EP_PREC_BF16X3, EP_PREC_FP16X3};
    std::vector<int4> tiles[4];
    for (int pi = 0; pi < 3; ++pi) {
        const bool f16 = precs[pi] == EP_PREC_FP16X3;
        std::vector<GemmArgs> st[4];
        for (int i = 0; i < n; ++i) {
            const BatchFieldInst &d = inst[i];
            const int si = (int)(std::find(bf->sizes.begin(), bf->sizes.end(), d.M) - bf->sizes.begin());
            const MatSet &s = bf->sets[si];
            const size_t nb = (size_t)d.M * d.M;
            uint32_t *fm = bf->fmax + 4 * i;
            GemmArgs a{};
            a.scale = (float)(4.0 / ((double)d.M * d.M) / ((double)d.bw * d.bh));
            a.wxs = (float)(M_PI / d.M);
            a.wys = (float)(M_PI / d.M * ((double)d.bw / d.bh));
            a.M = a.N = a.K = d.M;
            a.skip = d.skip;
            auto setA = [&](GemmArgs &g, int z, int m) {
                g.A[z] = s.f32[m];
                g.Ah[z] = f16 ? s.hi16[m] : s.hi[m];
                g.Al[z] = f16 ? s.lo16[m] : s.lo[m];
            };
            GemmArgs g1 = a;
            setA(g1, 0, MAT_CY); g1.B[0] = d.map_field; g1.C[0] = d.tA;
            g1.bmax = d.map_max; g1.cmax = f16 ? fm + 1 : nullptr; g1.bbias = d.map_mean;
            GemmArgs g2 = a;
            setA(g2, 0, MAT_CX); g2.B[0] = d.tA; g2.C[0] = d.tB; g2.C2[0] = d.tB + nb;
            g2.bmax = fm + 1; g2.cmax = f16 ? fm + 2 : nullptr;
            GemmArgs g3 = a;
            setA(g3, 0, MAT_CYT); g3.B[0] = d.tB; g3.C[0] = d.tA;
            setA(g3, 1, MAT_SYT); g3.B[1] = d.tB + nb; g3.C[1] = d.tA + nb;
            g3.bmax = fm + 2; g3.cmax = f16 ? fm + 3 : nullptr;
            GemmArgs g4 = a;
            setA(g4, 0, MAT_SXT); g4.B[0] = d.tA; g4.C[0] = d.ex;
            setA(g4, 1, MAT_CXT); g4.B[1] = d.tA + nb; g4.C[1] = d.ey;
            g4.bmax = fm + 3; g4.cmax = nullptr;
            st[0].push_back(g1);
            st[1].push_back(g2);
            st[2].push_back(g3);
            st[3].push_back(g4);
            if (pi == 0) {
                const int nt = (d.M + BM - 1) / BM;
                for (int stage = 0; stage < 4; ++stage)
                    for (int z = 0; z < (stage >= 2 ? 2 : 1); ++z)
                        for (int mt = 0; mt < nt; ++mt)
                            for (int ntl = 0; ntl < nt; ++ntl) tiles[stage].push_back(make_int4(i, mt * BM, ntl * BN, z));
            }
        }
        for (int stage = 0; stage < 4; ++stage) {
            if (hipMalloc(&bf->d_args[pi][stage], sizeof(GemmArgs) * std::max<size_t>(1, st[stage].size())) != hipSuccess ||
                hipMemcpy(bf->d_args[pi][stage], st[stage].data(), sizeof(GemmArgs) * st[stage].size(),
                          hipMemcpyHostToDevice) != hipSuccess)
                return fail(EP_ENOMEM, "batch field: argument upload failed");
        }
    }
    for (int stage = 0; stage < 4; ++stage) {
        bf->n_tiles[stage] = (int)tiles[stage].size();
        if (hipMalloc(&bf->d_tiles[stage], sizeof(int4) * std::max<size_t>(1, tiles[stage].size())) != hipSuccess ||
            hipMemcpy(bf->d_tiles[stage], tiles[stage].data(), sizeof(int4) * tiles[stage].size(),
                      hipMemcpyHostToDevice) != hipSuccess)
            return fail(EP_ENOMEM, "batch field: tile upload failed");
    }
    return EP_OK;
}

void batch_field_free(BatchField *bf)
{
    if (!bf) return;
    for (MatSet &s : bf->sets)
        for (int m = 0; m < MAT_N; ++m) {
            void *p[] = {s.f32[m], s.hi[m], s.lo[m], s.hi16[m], s.lo16[m]};
            for (void *q : p)
                if (q) (void)hipFree(q);
        }
    for (int pi = 0; pi < 3; ++pi)
        for (int st = 0; st < 4; ++st)
            if (bf->d_args[pi][st]) (void)hipFree(bf->d_args[pi][st]);
    for (int st = 0; st < 4; ++st)
        if (bf->d_tiles[st]) (void)hipFree(bf->d_tiles[st]);
    if (bf->fmax) (void)hipFree(bf->fmax);
    delete bf;
}

void launch_field_batch(BatchField *bf, int precision, hipStream_t s)
{
    const int pi = precision == EP_PREC_BF16X3 ? 1 : precision == EP_PREC_FP16X3 ? 2 : 0;
    if (precision == EP_PREC_FP16X3)
        hipLaunchKernelGGL(k_clear_max_b, dim3((bf->n_inst + 63) / 64), dim3(64), 0, s, bf->fmax, bf->n_inst);
    for (int stage = 0; stage < 4; ++stage) {
        const dim3 grid(bf->n_tiles[stage]);
        const GemmArgs *args = bf->d_args[pi][stage];
        const int4 *tiles = bf->d_tiles[stage];
        if (precision == EP_PREC_FP32) {
            if (stage == 1)
                hipLaunchKernelGGL(k_gemm_f32_b<1>, grid, dim3(256), 0, s, args, tiles);
            else
                hipLaunchKernelGGL(k_gemm_f32_b<0>, grid, dim3(256), 0, s, args, tiles);
        } else if (precision == EP_PREC_BF16X3) {
            if (stage == 1)
                hipLaunchKernelGGL((k_gemm_split_b<1, false>), grid, dim3(256), 0, s, args, tiles);
            else
                hipLaunchKernelGGL((k_gemm_split_b<0, false>), grid, dim3(256), 0, s, args, tiles);
        } else {
            if (stage == 1)
                hipLaunchKernelGGL((k_gemm_split_b<1, true>), grid, dim3(256), 0, s, args, tiles);
            else
                hipLaunchKernelGGL((k_gemm_split_b<0, true>), grid, dim3(256), 0, s, args, tiles);
        }
    }
}

}  // namespace ep
