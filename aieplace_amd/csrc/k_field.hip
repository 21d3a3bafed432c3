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
// Precision forms: FP32 (v_mfma_f32_32x32x2_f32) and split BF16 (each operand
// x = hi + lo in BF16, product hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_bf16
// with FP32 accumulation).  The constant matrices are split once at set-up;
// the data operand is split while it is staged into LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "ep_ctx.h"

namespace ep {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int BM = 64, BN = 64;

struct GemmArgs {
    const float *A[2];          // FP32 constant operand [M][K]
    const uint16_t *Ah[2];      // split constant operand, hi plane (BF16 bits)
    const uint16_t *Al[2];      //                          lo plane
    const float *B[2];          // data operand [N][K], FP32
    float *C[2];                // [M][N]
    float *C2[2];               // second output of the spectral epilogue
    int M, N, K;
    float scale;  // spectral: 4/(Mx My) / bin area
    float wxs;    // wx = wxs * u
    float wys;    // wy = wys * v
};

__device__ inline void store_out(const GemmArgs &a, int z, int epi, int row, int col, float v)
{
    const size_t o = (size_t)row * a.N + col;
    if (epi == 0) {
        a.C[z][o] = v;
    } else {
        float val = v * a.scale;
        if (row == 0) val *= 0.5f;
        if (col == 0) val *= 0.5f;
        const float wx = a.wxs * (float)row, wy = a.wys * (float)col;
        const float p = (row == 0 && col == 0) ? 0.0f : val / (wx * wx + wy * wy);
        a.C[z][o] = p * wx;
        a.C2[z][o] = p * wy;
    }
}

// ---------------------------------------------------------------- FP32

constexpr int BK32 = 16;

// 64x64 block tile, 4 waves of 32x32; global tiles [64][16] of A and Bt are
// staged transposed into LDS ([k][m], [k][n]) for the 32x32x2 operand layout.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f32(GemmArgs a)
{
    const int z = blockIdx.z;
    const float *__restrict__ A = a.A[z];
    const float *__restrict__ B = a.B[z];
    const int M = a.M, N = a.N, K = a.K;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    __shared__ float As[BK32][BM + 4];
    __shared__ float Bs[BK32][BN + 4];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    const int lr = t >> 2, lk = (t & 3) * 4;
    const bool a_ok = (m0 + lr) < M, b_ok = (n0 + lr) < N;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

    float4 av = make_float4(0, 0, 0, 0), bv = make_float4(0, 0, 0, 0);
    if (a_ok) av = *reinterpret_cast<const float4 *>(&A[(size_t)(m0 + lr) * K + lk]);
    if (b_ok) bv = *reinterpret_cast<const float4 *>(&B[(size_t)(n0 + lr) * K + lk]);
    for (int k0 = 0; k0 < K; k0 += BK32) {
        __syncthreads();
        As[lk + 0][lr] = av.x; As[lk + 1][lr] = av.y; As[lk + 2][lr] = av.z; As[lk + 3][lr] = av.w;
        Bs[lk + 0][lr] = bv.x; Bs[lk + 1][lr] = bv.y; Bs[lk + 2][lr] = bv.z; Bs[lk + 3][lr] = bv.w;
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

// ---------------------------------------------------------------- split BF16

constexpr int BK16 = 32;
constexpr int LDK = BK16 + 8;  // padded LDS row (bf16 elements)

__device__ inline uint16_t bf16_rne(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// 64x64 block tile, 4 waves of 32x32, K step 32 (two 32x32x16 MFMA sub-steps,
// three products each).  LDS rows are [m or n][k] with k contiguous, so each
// lane reads its 8 consecutive k values with one 16-byte load.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_bf16x3(GemmArgs a)
{
    const int z = blockIdx.z;
    const uint16_t *__restrict__ Ah = a.Ah[z];
    const uint16_t *__restrict__ Al = a.Al[z];
    const float *__restrict__ B = a.B[z];
    const int M = a.M, N = a.N, K = a.K;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    __shared__ __attribute__((aligned(16))) uint16_t sAh[BM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sAl[BM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBh[BN][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBl[BN][LDK];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    const int lr = t >> 2, lk = (t & 3) * 8;  // row and first k of this thread's 8-element chunk
    const bool a_ok = (m0 + lr) < M, b_ok = (n0 + lr) < N;
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
            for (int e = 0; e < 8; ++e) {
                h[e] = bf16_rne(v[e]);
                l[e] = bf16_rne(v[e] - bf16_to_f(h[e]));
            }
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
            const bf16x8 xh = *reinterpret_cast<const bf16x8 *>(&sAh[wm + (lane & 31)][kk]);
            const bf16x8 xl = *reinterpret_cast<const bf16x8 *>(&sAl[wm + (lane & 31)][kk]);
            const bf16x8 yh = *reinterpret_cast<const bf16x8 *>(&sBh[wn + (lane & 31)][kk]);
            const bf16x8 yl = *reinterpret_cast<const bf16x8 *>(&sBl[wn + (lane & 31)][kk]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, yh, acc, 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const int col = n0 + wn + (lane & 31);
        if (row < M && col < N) store_out(a, z, EPI, row, col, acc[i]);
    }
}

void gemm(const GemmArgs &a, int batch, int epi, int precision, hipStream_t s)
{
    dim3 grid((a.N + BN - 1) / BN, (a.M + BM - 1) / BM, batch);
    if (precision == EP_PREC_BF16X3) {
        if (epi == 0)
            hipLaunchKernelGGL(k_gemm_bf16x3<0>, grid, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(k_gemm_bf16x3<1>, grid, dim3(256), 0, s, a);
    } else {
        if (epi == 0)
            hipLaunchKernelGGL(k_gemm_f32<0>, grid, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(k_gemm_f32<1>, grid, dim3(256), 0, s, a);
    }
}

// T[r][c] = cos|sin(pi * f * (x + 1/2) / M) with (f, x) = (r, c) or, transposed, (c, r).
std::vector<float> trig(int M, bool sine, bool transpose)
{
    std::vector<float> t((size_t)M * M);
    for (int f = 0; f < M; ++f)
        for (int x = 0; x < M; ++x) {
            const double ang = M_PI * (double)f * ((double)x + 0.5) / (double)M;
            const float v = (float)(sine ? std::sin(ang) : std::cos(ang));
            t[transpose ? (size_t)x * M + f : (size_t)f * M + x] = v;
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

}  // namespace

// Matrix slots (FP32 and split copies share the index).
enum { MAT_CY = 0, MAT_CX, MAT_CYT, MAT_SYT, MAT_SXT, MAT_CXT, MAT_N };

struct FieldMats {
    float *f32[MAT_N];
    uint16_t *hi[MAT_N];
    uint16_t *lo[MAT_N];
};

ep_status field_init(Ctx &c)
{
    FieldMats *fm = new (std::nothrow) FieldMats();
    if (!fm) return fail(EP_ENOMEM, "field_init: host allocation failed");
    c.split_mats = fm;
    const int Ms[MAT_N] = {c.My, c.Mx, c.My, c.My, c.Mx, c.Mx};
    const bool sine[MAT_N] = {false, false, false, true, true, false};
    const bool tr[MAT_N] = {false, false, true, true, true, true};
    for (int m = 0; m < MAT_N; ++m) {
        std::vector<float> f = trig(Ms[m], sine[m], tr[m]);
        std::vector<uint16_t> h(f.size()), l(f.size());
        for (size_t k = 0; k < f.size(); ++k) {
            h[k] = bf16_rne_host(f[k]);
            l[k] = bf16_rne_host(f[k] - bf16_to_f_host(h[k]));
        }
        if (hipMalloc(&fm->f32[m], f.size() * sizeof(float)) != hipSuccess ||
            hipMalloc(&fm->hi[m], h.size() * sizeof(uint16_t)) != hipSuccess ||
            hipMalloc(&fm->lo[m], l.size() * sizeof(uint16_t)) != hipSuccess)
            return fail(EP_ENOMEM, "field_init: hipMalloc failed");
        if (hipMemcpy(fm->f32[m], f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(fm->hi[m], h.data(), h.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(fm->lo[m], l.data(), l.size() * 2, hipMemcpyHostToDevice) != hipSuccess)
            return fail(EP_EDEVICE, "field_init: upload failed");
    }
    return EP_OK;
}

void field_free(Ctx &c)
{
    FieldMats *fm = static_cast<FieldMats *>(c.split_mats);
    if (!fm) return;
    for (int m = 0; m < MAT_N; ++m) {
        if (fm->f32[m]) (void)hipFree(fm->f32[m]);
        if (fm->hi[m]) (void)hipFree(fm->hi[m]);
        if (fm->lo[m]) (void)hipFree(fm->lo[m]);
    }
    delete fm;
    c.split_mats = nullptr;
}

void launch_field(Ctx &c, int precision, hipStream_t s)
{
    FieldMats *fm = static_cast<FieldMats *>(c.split_mats);
    const size_t nb = (size_t)c.Mx * c.My;
    GemmArgs a{};
    a.scale = (float)(4.0 / ((double)c.Mx * c.My) / ((double)c.bw * c.bh));
    a.wxs = (float)(M_PI / c.Mx);
    a.wys = (float)(M_PI / c.My * ((double)c.bw / c.bh));
    auto setA = [&](int z, int m) {
        a.A[z] = fm->f32[m];
        a.Ah[z] = fm->hi[m];
        a.Al[z] = fm->lo[m];
    };
    // G1: T1t[v][i] = Cy * rho^T
    a.M = c.My; a.N = c.Mx; a.K = c.My;
    setA(0, MAT_CY); a.B[0] = c.map_field; a.C[0] = c.tA;
    gemm(a, 1, 0, precision, s);
    // G2: A[u][v] = Cx * T1t^T, spectral -> Ehx (tB), Ehy (tB + nb)
    a.M = c.Mx; a.N = c.My; a.K = c.Mx;
    setA(0, MAT_CX); a.B[0] = c.tA; a.C[0] = c.tB; a.C2[0] = c.tB + nb;
    gemm(a, 1, 1, precision, s);
    // G3: R1t[j][u] = CyT * Ehx^T, R2t = SyT * Ehy^T
    a.M = c.My; a.N = c.Mx; a.K = c.My;
    setA(0, MAT_CYT); a.B[0] = c.tB; a.C[0] = c.tA;
    setA(1, MAT_SYT); a.B[1] = c.tB + nb; a.C[1] = c.tA + nb;
    gemm(a, 2, 0, precision, s);
    // G4: Ex[i][j] = SxT * R1t^T, Ey = CxT * R2t^T
    a.M = c.Mx; a.N = c.My; a.K = c.Mx;
    setA(0, MAT_SXT); a.B[0] = c.tA; a.C[0] = c.ex;
    setA(1, MAT_CXT); a.B[1] = c.tA + nb; a.C[1] = c.ey;
    gemm(a, 2, 0, precision, s);
}

}  // namespace ep
