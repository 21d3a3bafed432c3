// k_field.hip -- Poisson field solve as dense transform-matrix products on MFMA.
//
// Reference: the separable DCT field of vck5000/host/src/AIEplace.cpp:752-934
// (compute_a_uv_DCT / compute_eField_DCT with DCT_naive/IDCT_naive/IDXST_naive,
// DCT.cpp:24-82) and the AIE DCT/IDCT/IDXST graphs (DCTGraph.h, IDCTGraph.h,
// IDXSTGraph.h), in the RePlAce scaling of python/AIEmath/fft_replace/fft.cpp:118-212.
// Each 1-D transform pass over all rows/columns is one GEMM with a cos/sin
// matrix; the spectral scaling is fused into the epilogue of the second GEMM.
//
// Map layout: rho[i*My + j], i = x bin, j = y bin.
//   T  = rho  * CyT          (CyT[j][v] = cos(pi v (j+1/2)/My))
//   A  = Cx   * T            (Cx[u][i]  = cos(pi u (i+1/2)/Mx)) -> spectral -> Ehx, Ehy
//   R1 = Ehx * Cy,  R2 = Ehy * Sy          (Cy[v][j], Sy[v][j] = sin(...))
//   Ex = SxT * R1,  Ey = CxT * R2          (SxT[i][u] = sin(pi u (i+1/2)/Mx))
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ep_ctx.h"

namespace ep {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int BM = 64, BN = 64, BK = 16;

struct GemmArgs {
    const float *A[2];
    const float *B[2];
    float *C[2];
    float *C2[2];
    int M, N, K;
    float scale;  // spectral: 4/(Mx My) / bin area
    float wxs;    // wx = wxs * u
    float wys;    // wy = wys * v
};

// C = A * B, row-major FP32, 64x64 block tile, 4 waves of 32x32 each on
// v_mfma_f32_32x32x2_f32.  M, N, K multiples of 16 (ragged 64-tiles masked).
// EPI = 1 applies the spectral step to the product and writes Ehx to C, Ehy to C2.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f32(GemmArgs a)
{
    const int z = blockIdx.z;
    const float *__restrict__ A = a.A[z];
    const float *__restrict__ B = a.B[z];
    const int M = a.M, N = a.N, K = a.K;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    __shared__ float As[BK][BM + 4];
    __shared__ float Bs[BK][BN + 4];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    const int ar = t >> 2, ak = (t & 3) * 4;
    const int bk = t >> 4, bn = (t & 15) * 4;
    const bool a_ok = (m0 + ar) < M, b_ok = (n0 + bn) < N;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

    float4 av = make_float4(0, 0, 0, 0), bv = make_float4(0, 0, 0, 0);
    if (a_ok) av = *reinterpret_cast<const float4 *>(&A[(size_t)(m0 + ar) * K + ak]);
    if (b_ok) bv = *reinterpret_cast<const float4 *>(&B[(size_t)bk * N + n0 + bn]);
    for (int k0 = 0; k0 < K; k0 += BK) {
        __syncthreads();
        As[ak + 0][ar] = av.x;
        As[ak + 1][ar] = av.y;
        As[ak + 2][ar] = av.z;
        As[ak + 3][ar] = av.w;
        *reinterpret_cast<float4 *>(&Bs[bk][bn]) = bv;
        __syncthreads();
        const int kn = k0 + BK;
        if (kn < K) {  // prefetch the next tile while the MFMAs run
            if (a_ok) av = *reinterpret_cast<const float4 *>(&A[(size_t)(m0 + ar) * K + kn + ak]);
            if (b_ok) bv = *reinterpret_cast<const float4 *>(&B[(size_t)(kn + bk) * N + n0 + bn]);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
            const float x = As[kk + (lane >> 5)][wm + (lane & 31)];
            const float y = Bs[kk + (lane >> 5)][wn + (lane & 31)];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x, y, acc, 0, 0, 0);
        }
    }
    float *__restrict__ C = a.C[z];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const int col = n0 + wn + (lane & 31);
        if (row >= M || col >= N) continue;
        const size_t o = (size_t)row * N + col;
        if (EPI == 0) {
            C[o] = acc[i];
        } else {
            float val = acc[i] * a.scale;
            if (row == 0) val *= 0.5f;
            if (col == 0) val *= 0.5f;
            const float wx = a.wxs * (float)row, wy = a.wys * (float)col;
            const float p = (row == 0 && col == 0) ? 0.0f : val / (wx * wx + wy * wy);
            C[o] = p * wx;
            a.C2[z][o] = p * wy;
        }
    }
}

void gemm(const GemmArgs &a, int batch, int epi, hipStream_t s)
{
    dim3 grid((a.N + BN - 1) / BN, (a.M + BM - 1) / BM, batch);
    if (epi == 0)
        hipLaunchKernelGGL(k_gemm_f32<0>, grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_gemm_f32<1>, grid, dim3(256), 0, s, a);
}

std::vector<float> trig(int M, bool sine, bool transpose)
{
    std::vector<float> t((size_t)M * M);
    for (int u = 0; u < M; ++u)
        for (int i = 0; i < M; ++i) {
            const double ang = M_PI * (double)u * ((double)i + 0.5) / (double)M;
            const float v = (float)(sine ? std::sin(ang) : std::cos(ang));
            t[transpose ? (size_t)i * M + u : (size_t)u * M + i] = v;
        }
    return t;
}

ep_status upload(float **dst, const std::vector<float> &v)
{
    if (hipMalloc(dst, v.size() * sizeof(float)) != hipSuccess)
        return fail(EP_ENOMEM, "field: hipMalloc of %zu floats failed", v.size());
    if (hipMemcpy(*dst, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return fail(EP_EDEVICE, "field: upload failed");
    return EP_OK;
}

}  // namespace

ep_status field_init(Ctx &c)
{
    ep_status st;
    if ((st = upload(&c.Cx, trig(c.Mx, false, false))) != EP_OK) return st;
    if ((st = upload(&c.Cx_T, trig(c.Mx, false, true))) != EP_OK) return st;
    if ((st = upload(&c.Sx_T, trig(c.Mx, true, true))) != EP_OK) return st;
    if ((st = upload(&c.Cy, trig(c.My, false, false))) != EP_OK) return st;
    if ((st = upload(&c.Cy_T, trig(c.My, false, true))) != EP_OK) return st;
    if ((st = upload(&c.Sy, trig(c.My, true, false))) != EP_OK) return st;
    return EP_OK;
}

void field_free(Ctx &c)
{
    float **m[] = {&c.Cx, &c.Cx_T, &c.Sx_T, &c.Cy, &c.Cy_T, &c.Sy};
    for (float **p : m)
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
}

void launch_field(Ctx &c, int precision, hipStream_t s)
{
    (void)precision;
    const size_t nb = (size_t)c.Mx * c.My;
    GemmArgs a{};
    a.scale = (float)(4.0 / ((double)c.Mx * c.My) / ((double)c.bw * c.bh));
    a.wxs = (float)(M_PI / c.Mx);
    a.wys = (float)(M_PI / c.My * ((double)c.bw / c.bh));
    // T = rho * CyT
    a.M = c.Mx; a.N = c.My; a.K = c.My;
    a.A[0] = c.map_field; a.B[0] = c.Cy_T; a.C[0] = c.tA;
    gemm(a, 1, 0, s);
    // A = Cx * T, spectral -> Ehx (tB), Ehy (tB + nb)
    a.M = c.Mx; a.N = c.My; a.K = c.Mx;
    a.A[0] = c.Cx; a.B[0] = c.tA; a.C[0] = c.tB; a.C2[0] = c.tB + nb;
    gemm(a, 1, 1, s);
    // R1 = Ehx * Cy, R2 = Ehy * Sy
    a.M = c.Mx; a.N = c.My; a.K = c.My;
    a.A[0] = c.tB; a.B[0] = c.Cy; a.C[0] = c.tA;
    a.A[1] = c.tB + nb; a.B[1] = c.Sy; a.C[1] = c.tA + nb;
    gemm(a, 2, 0, s);
    // Ex = SxT * R1, Ey = CxT * R2
    a.M = c.Mx; a.N = c.My; a.K = c.Mx;
    a.A[0] = c.Sx_T; a.B[0] = c.tA; a.C[0] = c.ex;
    a.A[1] = c.Cx_T; a.B[1] = c.tA + nb; a.C[1] = c.ey;
    gemm(a, 2, 0, s);
}

}  // namespace ep
