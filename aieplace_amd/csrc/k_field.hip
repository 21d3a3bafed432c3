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
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

constexpr float kF16Top = 14.0f;  // scaled data max lies in [2^14, 2^15)

struct GemmArgs {
    const float *A[2];          // FP32 even half of the constant operand [M/2][K/2]
    const float *Ao[2];         //      odd half
    const uint16_t *Ah[2];      // split even half, hi plane (BF16 or FP16 bits)
    const uint16_t *Al[2];      //                  lo plane
    const uint16_t *Aoh[2];     // split odd half, hi plane
    const uint16_t *Aol[2];     //                 lo plane
    float sgn[2];               // inverse products: +1 cosine, -1 sine
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
    const int32_t *skip;        // nonzero: the product is skipped (finished run or instance)
    int c_pair;                 // 1: C[z] is one half of an interleaved pair map (element o at 2 o)
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
        a.C[z][a.c_pair ? 2 * o : o] = v;
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

// ---------------------------------------------------------------- symmetric products
//
// Every transform matrix is even or odd about its centre:
//   forward  T[f][M-1-x]  = (-1)^f T[f][x]                 (cosine rows)
//   inverse  T'[M-1-x][f] = s (-1)^f T'[x][f],  s = +1 cosine, -1 sine
// so each M x K product splits into two (M/2) x (K/2) products over the even
// and odd halves E and O (half the flops):
//   forward  B+[n][x] = b[x] + b[K-1-x],  B-[n][x] = b[x] - b[K-1-x]   (x < K/2)
//            C[2r] = (Ae B+)[r],  C[2r+1] = (Ao B-)[r]
//   inverse  Be[n][f] = b[2f],  Bo[n][f] = b[2f+1]
//            E = Ae Be, O = Ao Bo;  C[x] = E + O,  C[M-1-x] = s (E - O)   (x < M/2)
// Ae/Ao are the even/odd rows (forward) or columns (inverse) of the left half
// of T, stored (M/2) x (K/2) row-major.  The forward butterfly runs while the
// data operand is staged into LDS; the inverse one in the epilogue.
//
// Block tile: 64 half-rows x 32 columns of both E and O.  Waves 0-1 hold E
// (half-rows 0-31 and 32-63), waves 2-3 hold O; each wave one 32x32 MFMA
// accumulator.
constexpr int SBM = 64, SBN = 32;

// Stores a wave's accumulator (times mul) through the symmetric epilogue;
// returns the largest |value| written.  xch: >= SBM * SBN floats of LDS that
// the caller no longer reads.
template <bool INV, int EPI>
__device__ inline float sym_store(const GemmArgs &a, int z, int h0, int n0, const f32x16 &acc,
                                  float mul, float *xch)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int eo = wv >> 1, wm = (wv & 1) * 32;
    const int Mh = a.M >> 1, col = n0 + (lane & 31);
    float m = 0.0f;
    if (!INV) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = h0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
            if (r < Mh && col < a.N) m = fmaxf(m, store_out(a, z, EPI, 2 * r + eo, col, acc[i] * mul));
        }
        return m;
    }
    __syncthreads();  // every wave is done with the staged operands
    if (eo) {
#pragma unroll
        for (int i = 0; i < 16; ++i)
            xch[(wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3)) * SBN + (lane & 31)] = acc[i];
    }
    __syncthreads();
    if (!eo) {
        const float sg = a.sgn[z];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int rl = wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3), r = h0 + rl;
            const float e = acc[i], o = xch[rl * SBN + (lane & 31)];
            if (r < Mh && col < a.N) {
                m = fmaxf(m, store_out(a, z, 0, r, col, (e + o) * mul));
                m = fmaxf(m, store_out(a, z, 0, a.M - 1 - r, col, sg * (e - o) * mul));
            }
        }
    }
    return m;
}

// ---------------------------------------------------------------- FP32
//
// SK half-k per step; the next step's operands are loaded while the MFMAs
// run.  The k order of the MFMA chain does not depend on SK, so every SK gives
// bit-identical results.  The single-instance products stage through two LDS
// stages in turn (2 sizeof(StageF32<SK>) bytes of dynamic LDS per block), the
// batched ones through one, between two barriers.

template <int SK>
struct StageF32 {
    float a[2][SK][SBM + 4];     // [E|O][k][half-row]
    float b[2][SK][SBN + 4];     // [B+|B- or Be|Bo][k][column]
};

// PP: two LDS stages in turn, as in the split form (sym_split_tile).
template <bool INV, int EPI, int SK, bool PP>
__device__ inline void sym_f32_tile(const GemmArgs &a, int z, int h0, int n0)
{
    constexpr int P = SK / 16;  // load pieces per thread and step
    extern __shared__ __attribute__((aligned(16))) float gemm_lds[];
    StageF32<SK> *St = reinterpret_cast<StageF32<SK> *>(gemm_lds);  // PP: St[0], St[1]
    const int N = a.N, K = a.K, Mh = a.M >> 1, Kh = K >> 1;
    const float *__restrict__ Ae = a.A[z];
    const float *__restrict__ Ao = a.Ao[z];
    const float *__restrict__ B = a.B[z];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int eo = wv >> 1, wm = (wv & 1) * 32;
    const int ar = t >> 2, ak = (t & 3) * 4;  // A: 64 rows x 16 k per piece, 4 per thread and half
    const int br = t >> 3, bk = (t & 7) * 2;  // B: 32 rows x 16 half-k per piece, 2 per thread and plane
    const bool a_ok = h0 + ar < Mh, b_ok = n0 + br < N;
    const float bias = (!INV && a.bbias && b_ok) ? *a.bbias : 0.0f;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    float4 ae[P], ao[P];
    float2 b0[P], b1[P];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int ka = k0 + q * 16 + ak, kb = k0 + q * 16 + bk;
            ae[q] = ao[q] = make_float4(0, 0, 0, 0);
            b0[q] = b1[q] = make_float2(0, 0);
            if (a_ok && ka < Kh) {
                ae[q] = *reinterpret_cast<const float4 *>(&Ae[(size_t)(h0 + ar) * Kh + ka]);
                ao[q] = *reinterpret_cast<const float4 *>(&Ao[(size_t)(h0 + ar) * Kh + ka]);
            }
            if (b_ok && kb < Kh) {
                const float *row = B + (size_t)(n0 + br) * K;
                if (INV) {
                    const float4 v = *reinterpret_cast<const float4 *>(&row[2 * kb]);
                    b0[q] = make_float2(v.x, v.z);  // b[2k], b[2k+2]
                    b1[q] = make_float2(v.y, v.w);  // b[2k+1], b[2k+3]
                } else {
                    b0[q] = *reinterpret_cast<const float2 *>(&row[kb]);          // b[k], b[k+1]
                    b1[q] = *reinterpret_cast<const float2 *>(&row[K - 2 - kb]);  // b[K-2-k], b[K-1-k]
                }
            }
        }
    };
    auto stage = [&](StageF32<SK> &S) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int ka = q * 16 + ak, kb = q * 16 + bk;
            S.a[0][ka + 0][ar] = ae[q].x; S.a[0][ka + 1][ar] = ae[q].y;
            S.a[0][ka + 2][ar] = ae[q].z; S.a[0][ka + 3][ar] = ae[q].w;
            S.a[1][ka + 0][ar] = ao[q].x; S.a[1][ka + 1][ar] = ao[q].y;
            S.a[1][ka + 2][ar] = ao[q].z; S.a[1][ka + 3][ar] = ao[q].w;
            if (INV) {
                S.b[0][kb][br] = b0[q].x; S.b[0][kb + 1][br] = b0[q].y;
                S.b[1][kb][br] = b1[q].x; S.b[1][kb + 1][br] = b1[q].y;
            } else {
                const float d0 = b0[q].x - bias, d1 = b0[q].y - bias;  // b[k], b[k+1]
                const float m0 = b1[q].y - bias, m1 = b1[q].x - bias;  // b[K-1-k], b[K-2-k]
                S.b[0][kb][br] = d0 + m0; S.b[0][kb + 1][br] = d1 + m1;
                S.b[1][kb][br] = d0 - m0; S.b[1][kb + 1][br] = d1 - m1;
            }
        }
    };
    auto mma = [&](const StageF32<SK> &S) {
#pragma unroll
        for (int kk = 0; kk < SK; kk += 2) {
            const float x = S.a[eo][kk + (lane >> 5)][wm + (lane & 31)];
            const float y = S.b[eo][kk + (lane >> 5)][lane & 31];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x, y, acc, 0, 0, 0);
        }
    };
    if (PP) {
        load(0);
        stage(St[0]);
        __syncthreads();
        if (SK < Kh) load(SK);
        int cur = 0;
        for (int k0 = 0; k0 < Kh; k0 += SK) {
            if (k0 + SK < Kh) stage(St[cur ^ 1]);
            if (k0 + 2 * SK < Kh) load(k0 + 2 * SK);
            mma(St[cur]);
            __syncthreads();
            cur ^= 1;
        }
    } else {
        load(0);
        for (int k0 = 0; k0 < Kh; k0 += SK) {
            __syncthreads();
            stage(St[0]);
            __syncthreads();
            if (k0 + SK < Kh) load(k0 + SK);  // next tiles in flight during the MFMAs
            mma(St[0]);
        }
    }
    sym_store<INV, EPI>(a, z, h0, n0, acc, 1.0f, &St[0].a[0][0][0]);
}

// The single-instance products run beside the WA branch's gathers, whose waves
// share the SIMDs; their waves take issue priority over them (the gathers have
// slack until the join).
template <bool INV, int EPI, int SK>
__global__ __launch_bounds__(256) void k_sym_f32(GemmArgs a)
{
    if (a.skip && *a.skip) return;
    __builtin_amdgcn_s_setprio(3);
    sym_f32_tile<INV, EPI, SK, true>(a, blockIdx.z, blockIdx.y * SBM, blockIdx.x * SBN);
}

// ---------------------------------------------------------------- split 16-bit

// SK half-k per step (SK / 16 sub-steps of three 32x32x16 MFMAs), staged in
// dynamic LDS (two sizeof(StageSplit<SK>) stages per block for the
// single-instance products, one for the batched ones); as in the FP32 form the
// result does not depend on SK.
template <int SK>
struct StageSplit {
    static constexpr int LD = SK + 8;  // padded LDS row (16-bit elements)
    uint16_t ah[2][SBM][LD], al[2][SBM][LD];  // [E|O][half-row][k], hi and lo planes
    uint16_t bh[2][SBN][LD], bl[2][SBN][LD];  // [B+|B- or Be|Bo][column][k]
};


template <bool F16>
__device__ inline f32x16 mfma16(const uint16_t *pa, const uint16_t *pb, f32x16 acc)
{
    if (F16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(*reinterpret_cast<const f16x8 *>(pa),
                                                     *reinterpret_cast<const f16x8 *>(pb), acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(*reinterpret_cast<const bf16x8 *>(pa),
                                                  *reinterpret_cast<const bf16x8 *>(pb), acc, 0, 0, 0);
}

// Splits two values into 16-bit hi and lo parts, x0 in the low half of each
// word.  Both conversions are the packed hardware ones (one instruction for two
// values, rounding to nearest even); for finite values the BF16 bits equal
// rounding the FP32 pattern, (u + 0x7FFF + bit 16 of u) >> 16, as the host's
// bf16_rne_host does for the constant matrices.
template <bool F16>
__device__ inline void split2(float x0, float x1, uint32_t &h, uint32_t &l)
{
    const f32x2 x = {x0, x1};
    if (F16) {
        const f16x2 hh = __builtin_convertvector(x, f16x2);
        h = __builtin_bit_cast(uint32_t, hh);
        l = __builtin_bit_cast(uint32_t, __builtin_convertvector(x - __builtin_convertvector(hh, f32x2), f16x2));
    } else {
        h = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
        const f32x2 hf = {__uint_as_float(h << 16), __uint_as_float(h & 0xFFFF0000u)};
        l = __builtin_bit_cast(uint32_t, __builtin_convertvector(x - hf, bf16x2));
    }
}

// Splits four values into hi and lo planes and stores them as 4 x 16 bits each.
template <bool F16>
__device__ inline void split_store4(const float (&v)[4], uint16_t *dh, uint16_t *dl)
{
    uint32_t h[2], l[2];
    split2<F16>(v[0], v[1], h[0], l[0]);
    split2<F16>(v[2], v[3], h[1], l[1]);
    *reinterpret_cast<uint2 *>(dh) = make_uint2(h[0], h[1]);
    *reinterpret_cast<uint2 *>(dl) = make_uint2(l[0], l[1]);
}

// PP (single-map products): two LDS stages in turn.  While a step's MFMAs read
// one stage, the same waves split the next step's data operand into the other
// and issue the loads of the step after it, so the MFMA pipe works while the
// VALU and LDS do; one barrier per step.  Without PP (batched products, two
// blocks per CU): one stage, loaded between two barriers.  The MFMA chain's k
// order is the same either way, so both give bit-identical products.
template <bool INV, int EPI, bool F16, int SK, bool PP>
__device__ inline void sym_split_tile(const GemmArgs &a, int z, int h0, int n0)
{
    constexpr int P16 = SK / 32;  // load pieces per thread and step
    extern __shared__ __attribute__((aligned(16))) float gemm_lds[];
    StageSplit<SK> *St = reinterpret_cast<StageSplit<SK> *>(gemm_lds);  // PP: St[0], St[1]
    const int N = a.N, K = a.K, Mh = a.M >> 1, Kh = K >> 1;
    const uint16_t *__restrict__ Aeh = a.Ah[z];
    const uint16_t *__restrict__ Ael = a.Al[z];
    const uint16_t *__restrict__ Aoh = a.Aoh[z];
    const uint16_t *__restrict__ Aol = a.Aol[z];
    const float *__restrict__ B = a.B[z];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int eo = wv >> 1, wm = (wv & 1) * 32;
    const int ar = t >> 2, ak = (t & 3) * 8;  // A: 64 rows x 32 half-k per piece, 8 per thread and plane
    const int br = t >> 3, bk = (t & 7) * 4;  // B: 32 rows x 32 half-k per piece, 4 per thread and plane
    const bool a_ok = h0 + ar < Mh, b_ok = n0 + br < N;
    // FP16: power-of-two scale of the data operand (exact; undone in the
    // epilogue).  The forward butterfly adds two values: one binade lower.
    float bscale = 1.0f;
    if (F16) {
        const float bm = __uint_as_float(*a.bmax);
        const int top = (int)kF16Top - (INV ? 0 : 1);
        if (bm > 0.0f && isfinite(bm)) bscale = ldexpf(1.0f, top - ilogbf(bm));
    }
    const float bias = (!INV && a.bbias && b_ok) ? *a.bbias : 0.0f;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    uint4 aeh[P16], ael[P16], aoh[P16], aol[P16];
    float4 b0[P16], b1[P16];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < P16; ++q) {
            const int ka = k0 + q * 32 + ak, kb = k0 + q * 32 + bk;
            aeh[q] = ael[q] = aoh[q] = aol[q] = make_uint4(0, 0, 0, 0);
            b0[q] = b1[q] = make_float4(0, 0, 0, 0);
            if (a_ok && ka < Kh) {
                const size_t o = (size_t)(h0 + ar) * Kh + ka;
                aeh[q] = *reinterpret_cast<const uint4 *>(&Aeh[o]);
                ael[q] = *reinterpret_cast<const uint4 *>(&Ael[o]);
                aoh[q] = *reinterpret_cast<const uint4 *>(&Aoh[o]);
                aol[q] = *reinterpret_cast<const uint4 *>(&Aol[o]);
            }
            if (b_ok && kb < Kh) {
                const float *row = B + (size_t)(n0 + br) * K;
                if (INV) {
                    const float4 v0 = *reinterpret_cast<const float4 *>(&row[2 * kb]);
                    const float4 v1 = *reinterpret_cast<const float4 *>(&row[2 * kb + 4]);
                    b0[q] = make_float4(v0.x, v0.z, v1.x, v1.z);  // b[2k .. 2k+6 even]
                    b1[q] = make_float4(v0.y, v0.w, v1.y, v1.w);  // odd
                } else {
                    b0[q] = *reinterpret_cast<const float4 *>(&row[kb]);          // b[k .. k+3]
                    b1[q] = *reinterpret_cast<const float4 *>(&row[K - 4 - kb]);  // b[K-4-k .. K-1-k]
                }
            }
        }
    };
    // the loaded registers -> one LDS stage (A planes as they are; the data
    // operand through the forward butterfly and the hi/lo split)
    auto stage = [&](StageSplit<SK> &S) {
#pragma unroll
        for (int q = 0; q < P16; ++q) {
            const int ka = q * 32 + ak, kb = q * 32 + bk;
            *reinterpret_cast<uint4 *>(&S.ah[0][ar][ka]) = aeh[q];
            *reinterpret_cast<uint4 *>(&S.al[0][ar][ka]) = ael[q];
            *reinterpret_cast<uint4 *>(&S.ah[1][ar][ka]) = aoh[q];
            *reinterpret_cast<uint4 *>(&S.al[1][ar][ka]) = aol[q];
            float p[4], m[4];
            if (INV) {
                p[0] = b0[q].x; p[1] = b0[q].y; p[2] = b0[q].z; p[3] = b0[q].w;
                m[0] = b1[q].x; m[1] = b1[q].y; m[2] = b1[q].z; m[3] = b1[q].w;
            } else {
                const float d[4] = {b0[q].x - bias, b0[q].y - bias, b0[q].z - bias, b0[q].w - bias};
                const float mi[4] = {b1[q].w - bias, b1[q].z - bias, b1[q].y - bias, b1[q].x - bias};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    p[j] = d[j] + mi[j];
                    m[j] = d[j] - mi[j];
                }
            }
            if (F16) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    p[j] *= bscale;
                    m[j] *= bscale;
                }
            }
            split_store4<F16>(p, &S.bh[0][br][kb], &S.bl[0][br][kb]);
            split_store4<F16>(m, &S.bh[1][br][kb], &S.bl[1][br][kb]);
        }
    };
    auto mma = [&](const StageSplit<SK> &S) {
#pragma unroll
        for (int ks = 0; ks < SK; ks += 16) {
            const int kk = ks + 8 * (lane >> 5);
            const uint16_t *xh = &S.ah[eo][wm + (lane & 31)][kk], *xl = &S.al[eo][wm + (lane & 31)][kk];
            const uint16_t *yh = &S.bh[eo][lane & 31][kk], *yl = &S.bl[eo][lane & 31][kk];
            acc = mfma16<F16>(xh, yh, acc);
            acc = mfma16<F16>(xh, yl, acc);
            acc = mfma16<F16>(xl, yh, acc);
        }
    };
    if (PP) {
        load(0);
        stage(St[0]);
        __syncthreads();
        if (SK < Kh) load(SK);
        int cur = 0;
        for (int k0 = 0; k0 < Kh; k0 += SK) {
            if (k0 + SK < Kh) stage(St[cur ^ 1]);          // the next step, beside this step's MFMAs
            if (k0 + 2 * SK < Kh) load(k0 + 2 * SK);       // the step after it
            mma(St[cur]);
            __syncthreads();
            cur ^= 1;
        }
    } else {
        load(0);
        for (int k0 = 0; k0 < Kh; k0 += SK) {
            __syncthreads();
            stage(St[0]);
            __syncthreads();
            if (k0 + SK < Kh) load(k0 + SK);
            mma(St[0]);
        }
    }
    const float m = sym_store<INV, EPI>(a, z, h0, n0, acc, 1.0f / bscale,
                                        reinterpret_cast<float *>(&St[0].ah[0][0][0]));
    if (F16 && a.cmax) track_max(a.cmax, m);
}

template <bool INV, int EPI, bool F16, int SK>
__global__ __launch_bounds__(256) void k_sym_split(GemmArgs a)
{
    if (a.skip && *a.skip) return;
    __builtin_amdgcn_s_setprio(3);  // as k_sym_f32
    sym_split_tile<INV, EPI, F16, SK, true>(a, blockIdx.z, blockIdx.y * SBM, blockIdx.x * SBN);
}

// Batched products: block b runs tile tiles[b] = (instance, h0, n0, z) with the
// instance's argument block; finished instances are skipped (uniform per block).
template <bool INV, int EPI, int SK>
__global__ __launch_bounds__(256) void k_sym_f32_b(const GemmArgs *__restrict__ args,
                                                   const int4 *__restrict__ tiles)
{
    const int4 t = tiles[blockIdx.x];
    const GemmArgs &a = args[t.x];
    if (a.skip && *a.skip) return;
    __builtin_amdgcn_s_setprio(3);  // as k_sym_f32 (the batch forks its WA gradient too)
    sym_f32_tile<INV, EPI, SK, false>(a, t.w, t.y, t.z);
}

template <bool INV, int EPI, bool F16, int SK>
__global__ __launch_bounds__(256) void k_sym_split_b(const GemmArgs *__restrict__ args,
                                                     const int4 *__restrict__ tiles)
{
    const int4 t = tiles[blockIdx.x];
    const GemmArgs &a = args[t.x];
    if (a.skip && *a.skip) return;
    __builtin_amdgcn_s_setprio(3);  // as k_sym_f32 (the batch forks its WA gradient too)
    sym_split_tile<INV, EPI, F16, SK, false>(a, t.w, t.y, t.z);
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
__global__ __launch_bounds__(256) void k_field_err(const float2 *__restrict__ exy,
                                                   const double *__restrict__ exd,
                                                   const double *__restrict__ eyd, int nb,
                                                   double *__restrict__ part)
{
    double v[4] = {0, 0, 0, 0};
    for (int b = blockIdx.x * 256 + threadIdx.x; b < nb; b += gridDim.x * 256) {
        const float2 e = exy[b];
        v[0] = fmax(v[0], fabs((double)e.x - exd[b]));
        v[1] = fmax(v[1], fabs(exd[b]));
        v[2] = fmax(v[2], fabs((double)e.y - eyd[b]));
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

// K steps of the products: long steps for a single map (one block per CU
// anyway), shorter ones when two maps run so that two blocks share a CU.
constexpr int kSkF32Single = 32, kSkF32Batch = 32;
constexpr int kSkSplitSingle = 64, kSkSplitBatch = 32;  // two stages each: 110 KB (one map), 56 KB

template <int SK>
void gemm_f32(const GemmArgs &a, dim3 grid, bool inv, int epi, hipStream_t s)
{
    const size_t lds = 2 * sizeof(StageF32<SK>);  // two stages in turn
    if (inv) hipLaunchKernelGGL((k_sym_f32<true, 0, SK>), grid, dim3(256), lds, s, a);
    else if (epi) hipLaunchKernelGGL((k_sym_f32<false, 1, SK>), grid, dim3(256), lds, s, a);
    else hipLaunchKernelGGL((k_sym_f32<false, 0, SK>), grid, dim3(256), lds, s, a);
}

template <bool F16, int SK>
void gemm_split(const GemmArgs &a, dim3 grid, bool inv, int epi, hipStream_t s)
{
    const size_t lds = 2 * sizeof(StageSplit<SK>);  // two stages in turn
    if (inv) hipLaunchKernelGGL((k_sym_split<true, 0, F16, SK>), grid, dim3(256), lds, s, a);
    else if (epi) hipLaunchKernelGGL((k_sym_split<false, 1, F16, SK>), grid, dim3(256), lds, s, a);
    else hipLaunchKernelGGL((k_sym_split<false, 0, F16, SK>), grid, dim3(256), lds, s, a);
}

// One product (or `batch` independent ones) of the field solve.
void gemm(const GemmArgs &a, int batch, bool inv, int epi, int precision, hipStream_t s)
{
    const dim3 grid((a.N + SBN - 1) / SBN, ((a.M >> 1) + SBM - 1) / SBM, batch);
    const bool single = batch == 1;
    if (precision == EP_PREC_FP32) {
        if (single) gemm_f32<kSkF32Single>(a, grid, inv, epi, s);
        else gemm_f32<kSkF32Batch>(a, grid, inv, epi, s);
    } else if (precision == EP_PREC_FP16X3) {
        if (single) gemm_split<true, kSkSplitSingle>(a, grid, inv, epi, s);
        else gemm_split<true, kSkSplitBatch>(a, grid, inv, epi, s);
    } else {
        if (single) gemm_split<false, kSkSplitSingle>(a, grid, inv, epi, s);
        else gemm_split<false, kSkSplitBatch>(a, grid, inv, epi, s);
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

// Even and odd halves, (M/2) x (M/2) row-major, of one transform matrix:
//   forward (T[f][x]):  E[r][x] = T[2r][x],  O[r][x] = T[2r+1][x]     (x < M/2)
//   inverse (T'[x][f]): E[x][r] = T'[x][2r], O[x][r] = T'[x][2r+1]    (x < M/2)
// with T[f][x] = cos|sin(pi f (x + 1/2) / M) and T' its transpose.
void halves(int M, bool sine, bool inverse, std::vector<double> &E, std::vector<double> &O)
{
    const int h = M / 2;
    E.assign((size_t)h * h, 0.0);
    O.assign((size_t)h * h, 0.0);
    for (int r = 0; r < h; ++r)
        for (int x = 0; x < h; ++x) {
            const double ae = M_PI * (2.0 * r) * (x + 0.5) / M, ao = M_PI * (2.0 * r + 1.0) * (x + 0.5) / M;
            const size_t o = inverse ? (size_t)x * h + r : (size_t)r * h + x;
            E[o] = sine ? std::sin(ae) : std::cos(ae);
            O[o] = sine ? std::sin(ao) : std::cos(ao);
        }
}

}  // namespace

// Matrix slots (all copies share the index).
enum { MAT_CY = 0, MAT_CX, MAT_CYT, MAT_SYT, MAT_SXT, MAT_CXT, MAT_N };

// Transform of each slot: sine?, inverse?, and the inverse products' mirror sign.
constexpr bool kSlotSine[MAT_N] = {false, false, false, true, true, false};
constexpr bool kSlotInv[MAT_N] = {false, false, true, true, true, true};
constexpr float kSlotSign[MAT_N] = {1.f, 1.f, 1.f, -1.f, -1.f, 1.f};

// Device copies of one slot's halves ([0] even, [1] odd) in every precision.
struct HalfMats {
    float *f32[2] = {};
    uint16_t *bh[2] = {}, *bl[2] = {};   // BF16 hi/lo
    uint16_t *fh[2] = {}, *fl[2] = {};   // FP16 hi/lo
};

static ep_status upload_halves(int M, int slot, HalfMats &hm)
{
    std::vector<double> d[2];
    halves(M, kSlotSine[slot], kSlotInv[slot], d[0], d[1]);
    for (int k = 0; k < 2; ++k) {
        const size_t n = d[k].size();
        std::vector<float> f(n);
        std::vector<uint16_t> h(n), l(n), h16(n), l16(n);
        for (size_t e = 0; e < n; ++e) {
            f[e] = (float)d[k][e];
            h[e] = bf16_rne_host(f[e]);
            l[e] = bf16_rne_host(f[e] - bf16_to_f_host(h[e]));
            h16[e] = f16_host(f[e]);
            l16[e] = f16_host(f[e] - f16_to_f_host(h16[e]));
        }
        if (hipMalloc(&hm.f32[k], n * 4) != hipSuccess || hipMalloc(&hm.bh[k], n * 2) != hipSuccess ||
            hipMalloc(&hm.bl[k], n * 2) != hipSuccess || hipMalloc(&hm.fh[k], n * 2) != hipSuccess ||
            hipMalloc(&hm.fl[k], n * 2) != hipSuccess)
            return fail(EP_ENOMEM, "field: hipMalloc of transform matrices failed");
        if (hipMemcpy(hm.f32[k], f.data(), n * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(hm.bh[k], h.data(), n * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(hm.bl[k], l.data(), n * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(hm.fh[k], h16.data(), n * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(hm.fl[k], l16.data(), n * 2, hipMemcpyHostToDevice) != hipSuccess)
            return fail(EP_EDEVICE, "field: upload of transform matrices failed");
    }
    return EP_OK;
}

static void free_halves(HalfMats &hm)
{
    for (int k = 0; k < 2; ++k) {
        void *p[] = {hm.f32[k], hm.bh[k], hm.bl[k], hm.fh[k], hm.fl[k]};
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
    hm = HalfMats();
}

// Points product slot z of `a` at matrix slot m in the given precision.
static void set_operand(GemmArgs &a, int z, const HalfMats &hm, int m, int precision)
{
    const bool f16 = precision == EP_PREC_FP16X3;
    a.A[z] = hm.f32[0];
    a.Ao[z] = hm.f32[1];
    a.Ah[z] = f16 ? hm.fh[0] : hm.bh[0];
    a.Al[z] = f16 ? hm.fl[0] : hm.bl[0];
    a.Aoh[z] = f16 ? hm.fh[1] : hm.bh[1];
    a.Aol[z] = f16 ? hm.fl[1] : hm.bl[1];
    a.sgn[z] = kSlotSign[m];
}

struct FieldMats {
    HalfMats half[MAT_N];                  // even/odd halves, FP32 and split planes
    double *f64[MAT_N];                    // full FP64 matrices of the reference (built on first use)
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
        ep_status st = upload_halves(M, m, fm->half[m]);
        if (st != EP_OK) return st;
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
        free_halves(fm->half[m]);
        if (fm->f64[m]) (void)hipFree(fm->f64[m]);
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
    a.skip = &c.d_scal->halt;  // a finished device-controlled run skips the solve
    auto setA = [&](int z, int m) { set_operand(a, z, fm->half[m], m, precision); };
    if (f16) {
        launch_map_range(c, s);
        hipLaunchKernelGGL(k_clear_max, dim3(1), dim3(64), 0, s, fm->fmax + 1, 3);
    }
    // G1: T1t[v][i] = Cy * rho^T
    a.M = c.My; a.N = c.Mx; a.K = c.My;
    setA(0, MAT_CY); a.B[0] = c.map_field; a.C[0] = c.tA;
    a.bmax = fm->fmax + 0; a.cmax = f16 ? fm->fmax + 1 : nullptr;
    // the DC mode carries no field, so subtracting any constant changes only the
    // dropped DC coefficient; subtracting the mean keeps the FP32 and 16-bit sums small
    a.bbias = c.map_mean;
    gemm(a, 1, false, 0, precision, s);
    a.bbias = nullptr;
    // G2: A[u][v] = Cx * T1t^T, spectral -> Ehx (tB), Ehy (tB + nb)
    a.M = c.Mx; a.N = c.My; a.K = c.Mx;
    setA(0, MAT_CX); a.B[0] = c.tA; a.C[0] = c.tB; a.C2[0] = c.tB + nb;
    a.bmax = fm->fmax + 1; a.cmax = f16 ? fm->fmax + 2 : nullptr;
    gemm(a, 1, false, 1, precision, s);
    // G3: R1t[j][u] = CyT * Ehx^T, R2t = SyT * Ehy^T
    a.M = c.My; a.N = c.Mx; a.K = c.My;
    setA(0, MAT_CYT); a.B[0] = c.tB; a.C[0] = c.tA;
    setA(1, MAT_SYT); a.B[1] = c.tB + nb; a.C[1] = c.tA + nb;
    a.bmax = fm->fmax + 2; a.cmax = f16 ? fm->fmax + 3 : nullptr;
    gemm(a, 2, true, 0, precision, s);
    // G4: Ex[i][j] = SxT * R1t^T, Ey = CxT * R2t^T
    a.M = c.Mx; a.N = c.My; a.K = c.Mx;
    setA(0, MAT_SXT); a.B[0] = c.tA; a.C[0] = reinterpret_cast<float *>(c.exy);
    setA(1, MAT_CXT); a.B[1] = c.tA + nb; a.C[1] = reinterpret_cast<float *>(c.exy) + 1;
    a.c_pair = 1;
    a.bmax = fm->fmax + 3; a.cmax = nullptr;
    gemm(a, 2, true, 0, precision, s);
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
    hipLaunchKernelGGL(k_field_err, dim3(kErrBlocks), dim3(256), 0, s, c.exy, fm->exd, fm->eyd,
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
    HalfMats half[MAT_N];
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
    std::vector<int> M_of;                   // grid side of each instance
    int tile_var[3] = {1, 1, 1};             // [precision] K-step variant (kSkBatchF32/kSkBatchSplit)
};

ep_status batch_field_init(BatchField *&out, const std::vector<BatchFieldInst> &inst)
{
    BatchField *bf = new (std::nothrow) BatchField();
    if (!bf) return fail(EP_ENOMEM, "batch field: host allocation failed");
    out = bf;
    bf->n_inst = (int)inst.size();
    for (const BatchFieldInst &d : inst) bf->M_of.push_back(d.M);
    for (const BatchFieldInst &d : inst)
        if (std::find(bf->sizes.begin(), bf->sizes.end(), d.M) == bf->sizes.end()) bf->sizes.push_back(d.M);
    bf->sets.resize(bf->sizes.size());
    for (size_t si = 0; si < bf->sizes.size(); ++si)
        for (int m = 0; m < MAT_N; ++m) {
            ep_status st = upload_halves(bf->sizes[si], m, bf->sets[si].half[m]);
            if (st != EP_OK) return st;
        }
    const int n = bf->n_inst;
    if (hipMalloc(&bf->fmax, sizeof(uint32_t) * 4 * std::max(1, n)) != hipSuccess ||
        hipMemset(bf->fmax, 0, sizeof(uint32_t) * 4 * std::max(1, n)) != hipSuccess)
        return fail(EP_ENOMEM, "batch field: hipMalloc failed");
    // argument blocks per precision and stage, and the tile lists
    const int precs[3] = {EP_PREC_FP32, EP_PREC_BF16X3, EP_PREC_FP16X3};
    std::vector<int4> tiles[4];
    for (int pi = 0; pi < 3; ++pi) {
        const int prec = precs[pi];
        const bool f16 = prec == EP_PREC_FP16X3;
        std::vector<GemmArgs> st[4];
        for (int i = 0; i < n; ++i) {
            const BatchFieldInst &d = inst[i];
            const int si = (int)(std::find(bf->sizes.begin(), bf->sizes.end(), d.M) - bf->sizes.begin());
            const MatSet &ms = bf->sets[si];
            const size_t nb = (size_t)d.M * d.M;
            uint32_t *fm = bf->fmax + 4 * i;
            GemmArgs a{};
            a.scale = (float)(4.0 / ((double)d.M * d.M) / ((double)d.bw * d.bh));
            a.wxs = (float)(M_PI / d.M);
            a.wys = (float)(M_PI / d.M * ((double)d.bw / d.bh));
            a.M = a.N = a.K = d.M;
            a.skip = d.skip;
            auto setA = [&](GemmArgs &g, int z, int m) { set_operand(g, z, ms.half[m], m, prec); };
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
            setA(g4, 0, MAT_SXT); g4.B[0] = d.tA; g4.C[0] = reinterpret_cast<float *>(d.exy);
            setA(g4, 1, MAT_CXT); g4.B[1] = d.tA + nb; g4.C[1] = reinterpret_cast<float *>(d.exy) + 1;
            g4.c_pair = 1;
            g4.bmax = fm + 3; g4.cmax = nullptr;
            st[0].push_back(g1);
            st[1].push_back(g2);
            st[2].push_back(g3);
            st[3].push_back(g4);
            if (pi == 0) {
                const int mt = ((d.M >> 1) + SBM - 1) / SBM, nt = (d.M + SBN - 1) / SBN;
                for (int stage = 0; stage < 4; ++stage)
                    for (int z = 0; z < (stage >= 2 ? 2 : 1); ++z)
                        for (int r = 0; r < mt; ++r)
                            for (int q = 0; q < nt; ++q) tiles[stage].push_back(make_int4(i, r * SBM, q * SBN, z));
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

ep_status batch_field_set_active(BatchField *bf, const std::vector<char> &active)
{
    std::vector<int4> tiles[4];
    for (int i = 0; i < bf->n_inst; ++i) {
        if (!active[i]) continue;
        const int M = bf->M_of[i];
        const int mt = ((M >> 1) + SBM - 1) / SBM, nt = (M + SBN - 1) / SBN;
        for (int stage = 0; stage < 4; ++stage)
            for (int z = 0; z < (stage >= 2 ? 2 : 1); ++z)
                for (int r = 0; r < mt; ++r)
                    for (int q = 0; q < nt; ++q) tiles[stage].push_back(make_int4(i, r * SBM, q * SBN, z));
    }
    for (int stage = 0; stage < 4; ++stage) {  // the lists only shrink: the buffers fit
        bf->n_tiles[stage] = (int)tiles[stage].size();
        if (!tiles[stage].empty() &&
            hipMemcpy(bf->d_tiles[stage], tiles[stage].data(), sizeof(int4) * tiles[stage].size(),
                      hipMemcpyHostToDevice) != hipSuccess)
            return fail(EP_EDEVICE, "batch field: tile upload failed");
    }
    return EP_OK;
}

void batch_field_free(BatchField *bf)
{
    if (!bf) return;
    for (MatSet &ms : bf->sets)
        for (int m = 0; m < MAT_N; ++m) free_halves(ms.half[m]);
    for (int pi = 0; pi < 3; ++pi)
        for (int st = 0; st < 4; ++st)
            if (bf->d_args[pi][st]) (void)hipFree(bf->d_args[pi][st]);
    for (int st = 0; st < 4; ++st)
        if (bf->d_tiles[st]) (void)hipFree(bf->d_tiles[st]);
    if (bf->fmax) (void)hipFree(bf->fmax);
    delete bf;
}

namespace {

// K-step variants of the batched products (every variant gives the same
// result; the fastest depends on the grid sizes and the number of tiles).
constexpr int kSkBatchF32[3] = {16, 32, 64};
constexpr int kSkBatchSplit[3] = {32, 64, 128};

int prec_index(int precision)
{
    return precision == EP_PREC_BF16X3 ? 1 : precision == EP_PREC_FP16X3 ? 2 : 0;
}

template <int SF, int SS>
void launch_stage_b(int precision, int stage, dim3 grid, const GemmArgs *args, const int4 *tiles,
                    hipStream_t s)
{
    const dim3 blk(256);
    const bool inv = stage >= 2, epi = stage == 1;
    if (precision == EP_PREC_FP32) {
        const size_t lds = sizeof(StageF32<SF>);
        if (inv) hipLaunchKernelGGL((k_sym_f32_b<true, 0, SF>), grid, blk, lds, s, args, tiles);
        else if (epi) hipLaunchKernelGGL((k_sym_f32_b<false, 1, SF>), grid, blk, lds, s, args, tiles);
        else hipLaunchKernelGGL((k_sym_f32_b<false, 0, SF>), grid, blk, lds, s, args, tiles);
        return;
    }
    const size_t lds = sizeof(StageSplit<SS>);
    if (precision == EP_PREC_BF16X3) {
        if (inv) hipLaunchKernelGGL((k_sym_split_b<true, 0, false, SS>), grid, blk, lds, s, args, tiles);
        else if (epi) hipLaunchKernelGGL((k_sym_split_b<false, 1, false, SS>), grid, blk, lds, s, args, tiles);
        else hipLaunchKernelGGL((k_sym_split_b<false, 0, false, SS>), grid, blk, lds, s, args, tiles);
    } else {
        if (inv) hipLaunchKernelGGL((k_sym_split_b<true, 0, true, SS>), grid, blk, lds, s, args, tiles);
        else if (epi) hipLaunchKernelGGL((k_sym_split_b<false, 1, true, SS>), grid, blk, lds, s, args, tiles);
        else hipLaunchKernelGGL((k_sym_split_b<false, 0, true, SS>), grid, blk, lds, s, args, tiles);
    }
}

}  // namespace

void launch_field_batch(BatchField *bf, int precision, hipStream_t s)
{
    const int pi = prec_index(precision);
    if (precision == EP_PREC_FP16X3)
        hipLaunchKernelGGL(k_clear_max_b, dim3((bf->n_inst + 63) / 64), dim3(64), 0, s, bf->fmax, bf->n_inst);
    for (int stage = 0; stage < 4; ++stage) {
        if (bf->n_tiles[stage] == 0) continue;
        const dim3 grid(bf->n_tiles[stage]);
        const GemmArgs *args = bf->d_args[pi][stage];
        const int4 *tiles = bf->d_tiles[stage];
        switch (bf->tile_var[pi]) {
        case 0: launch_stage_b<kSkBatchF32[0], kSkBatchSplit[0]>(precision, stage, grid, args, tiles, s); break;
        case 1: launch_stage_b<kSkBatchF32[1], kSkBatchSplit[1]>(precision, stage, grid, args, tiles, s); break;
        default: launch_stage_b<kSkBatchF32[2], kSkBatchSplit[2]>(precision, stage, grid, args, tiles, s); break;
        }
    }
}

ep_status batch_field_autotune(BatchField *bf, int reps, hipStream_t s)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(EP_EDEVICE, "batch field autotune: hipEventCreate failed");
    }
    const int precs[3] = {EP_PREC_FP32, EP_PREC_BF16X3, EP_PREC_FP16X3};
    ep_status st = EP_OK;
    for (int pi = 0; pi < 3 && st == EP_OK; ++pi) {
        float best = 0.f;
        int best_v = 1;
        for (int v = 0; v < 3; ++v) {
            bf->tile_var[pi] = v;
            launch_field_batch(bf, precs[pi], s);  // first launch of this variant: untimed
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < reps; ++r) launch_field_batch(bf, precs[pi], s);
            (void)hipEventRecord(e1, s);
            hipError_t e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipGetLastError();
            if (e != hipSuccess) {
                st = fail(EP_EDEVICE, "batch field autotune: %s", hipGetErrorString(e));
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (v == 0 || ms < best) {
                best = ms;
                best_v = v;
            }
        }
        bf->tile_var[pi] = best_v;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

int batch_field_k_step(const BatchField *bf, int precision)
{
    const int v = bf->tile_var[prec_index(precision)];
    return precision == EP_PREC_FP32 ? kSkBatchF32[v] : kSkBatchSplit[v];
}

// ---------------------------------------------------------------- standalone transforms
//
// One reference-convention transform on the product path (ep_transform): the
// transformed axis is staged contiguous, IDCT/IDXST rows get x_0/2, IDXST rows
// are reversed (IDXST(x)_k = (-1)^k IDCT(x_0, x_{n-1}, ..., x_1)_k, DCT.cpp:64-82),
// and one product with the forward or inverse cosine matrix transforms every row.

namespace {

constexpr int kTrBlock = 256;

// X[r][k] = rowsign(r) * h(k) * src[r][k'] over R rows of length n, where
// k' = n - k for reversed rows (k > 0), h(0) = 1/2 when `halve`, and rowsign(r)
// = (-1)^r when `altsign` (the previous pass was an IDXST).  `tr`: the source is
// the n x R map whose columns are the rows (src[k'][r]).  Tracks max |X| in maxbits.
__global__ __launch_bounds__(kTrBlock) void k_tr_prep(const float *__restrict__ src, float *__restrict__ X,
                                                      int R, int n, int tr, int reverse, int halve, int altsign,
                                                      uint32_t *__restrict__ maxbits)
{
    float m = 0.0f;
    const int64_t total = (int64_t)R * n;
    for (int64_t e = (int64_t)blockIdx.x * kTrBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kTrBlock) {
        const int r = (int)(e / n), k = (int)(e - (int64_t)r * n);
        const int kk = (reverse && k > 0) ? n - k : k;
        float v = tr ? src[(int64_t)kk * R + r] : src[(int64_t)r * n + kk];
        if (halve && k == 0) v *= 0.5f;
        if (altsign && (r & 1)) v = -v;
        X[e] = v;
        m = fmaxf(m, fabsf(v));
    }
    track_max(maxbits, m);
}

// Product output C[f][r] (n x R) -> Y: sign (-1)^f when `altsign`; Y[r][f] when
// `tr`, else in the same layout.
__global__ __launch_bounds__(kTrBlock) void k_tr_post(const float *__restrict__ C, float *__restrict__ Y, int R,
                                                      int n, int altsign, int tr)
{
    const int64_t total = (int64_t)R * n;
    for (int64_t e = (int64_t)blockIdx.x * kTrBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kTrBlock) {
        const int f = (int)(e / R), r = (int)(e - (int64_t)f * R);
        float v = C[e];
        if (altsign && (f & 1)) v = -v;
        Y[tr ? (int64_t)r * n + f : e] = v;
    }
}

__global__ void k_tr_zero(uint32_t *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

int tr_grid(int64_t n)
{
    const int64_t g = (n + kTrBlock - 1) / kTrBlock;
    return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 4096);
}

}  // namespace

ep_status run_transform(int row_kind, int col_kind, int nx, int ny, const float *in, float *out, int precision)
{
    const size_t nb = (size_t)nx * ny;
    float *d_in = nullptr, *d_x = nullptr, *d_c = nullptr, *d_out = nullptr;
    uint32_t *d_max = nullptr;
    HalfMats mats[2];
    ep_status st = EP_OK;
    hipStream_t s = nullptr;
    auto cleanup = [&]() {
        void *p[] = {d_in, d_x, d_c, d_out, d_max};
        for (void *q : p)
            if (q) (void)hipFree(q);
        free_halves(mats[0]);
        free_halves(mats[1]);
    };
    if (hipMalloc(&d_in, nb * 4) != hipSuccess || hipMalloc(&d_x, nb * 4) != hipSuccess ||
        hipMalloc(&d_c, nb * 4) != hipSuccess || hipMalloc(&d_out, nb * 4) != hipSuccess ||
        hipMalloc(&d_max, 4 * sizeof(uint32_t)) != hipSuccess) {
        cleanup();
        return fail(EP_ENOMEM, "ep_transform: hipMalloc failed");
    }
    if (hipMemcpy(d_in, in, nb * 4, hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return fail(EP_EDEVICE, "ep_transform: upload failed");
    }
    hipLaunchKernelGGL(k_tr_zero, dim3(1), dim3(64), 0, s, d_max, 4);
    // One pass: rows of X (R x n, staged from src) through the product into C (n x R).
    auto pass = [&](int kind, int n, int R, const float *src, int tr, int altsign, HalfMats &hm,
                    uint32_t *bmax, uint32_t *cmax) -> ep_status {
        const int slot = kind == EP_TR_DCT ? MAT_CX : MAT_CXT;  // forward / inverse cosine
        ep_status e = upload_halves(n, slot, hm);
        if (e != EP_OK) return e;
        hipLaunchKernelGGL(k_tr_prep, dim3(tr_grid((int64_t)R * n)), dim3(kTrBlock), 0, s, src, d_x, R, n, tr,
                           kind == EP_TR_IDXST ? 1 : 0, kind == EP_TR_DCT ? 0 : 1, altsign, bmax);
        GemmArgs a{};
        set_operand(a, 0, hm, slot, precision);
        a.M = n;
        a.N = R;
        a.K = n;
        a.B[0] = d_x;
        a.C[0] = d_c;
        a.bmax = bmax;
        a.cmax = cmax;
        gemm(a, 1, kind != EP_TR_DCT, 0, precision, s);
        return EP_OK;
    };
    if (row_kind != EP_TR_NONE && col_kind != EP_TR_NONE) {
        // rows of in (nx rows of length ny) -> C1[v][i]; then the rows of C1
        // (ny rows of length nx) -> C2[u][v] = out
        if ((st = pass(row_kind, ny, nx, d_in, 0, 0, mats[0], d_max + 0, d_max + 1)) != EP_OK) {
            cleanup();
            return st;
        }
        if ((st = pass(col_kind, nx, ny, d_c, 0, row_kind == EP_TR_IDXST, mats[1], d_max + 2, d_max + 3)) !=
            EP_OK) {
            cleanup();
            return st;
        }
        hipLaunchKernelGGL(k_tr_post, dim3(tr_grid((int64_t)nb)), dim3(kTrBlock), 0, s, d_c, d_out, ny, nx,
                           col_kind == EP_TR_IDXST ? 1 : 0, 0);
    } else if (row_kind != EP_TR_NONE) {
        if ((st = pass(row_kind, ny, nx, d_in, 0, 0, mats[0], d_max + 0, d_max + 1)) != EP_OK) {
            cleanup();
            return st;
        }
        hipLaunchKernelGGL(k_tr_post, dim3(tr_grid((int64_t)nb)), dim3(kTrBlock), 0, s, d_c, d_out, nx, ny,
                           row_kind == EP_TR_IDXST ? 1 : 0, 1);
    } else {
        // columns of in: staged transposed (ny rows of length nx) -> C[u][j] = out
        if ((st = pass(col_kind, nx, ny, d_in, 1, 0, mats[0], d_max + 0, d_max + 1)) != EP_OK) {
            cleanup();
            return st;
        }
        hipLaunchKernelGGL(k_tr_post, dim3(tr_grid((int64_t)nb)), dim3(kTrBlock), 0, s, d_c, d_out, ny, nx,
                           col_kind == EP_TR_IDXST ? 1 : 0, 0);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out, d_out, nb * 4, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return fail(EP_EDEVICE, "ep_transform: %s", hipGetErrorString(e));
    return EP_OK;
}

}  // namespace ep
