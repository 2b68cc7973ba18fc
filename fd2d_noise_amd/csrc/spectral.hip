// spectral.hip -- frequency-domain slab products of the engine on the FP32 matrix cores.
//
// Both contractions are tall-skinny products over the P grid nodes, with the complex phase
// factors of a slab of time samples split into real columns c = 2k (real part) and c = 2k + 1:
//   flush:  acc[k][x] += sum_j samp[j][x] * e[j][k]               (the DFT of a wavefield,
//           code/run_forward1_green.m:176-183, run_noise_adjoint.m:335-346)
//   synth:  out[j][x]  = sum_c et[c][j] * a[c][x]                  (the distributed source of the
//           sampled steps, run_forward2_correlation.m:195-216: Re(A e^{iwt}) = Ar*cr - Ai*ci)
// Samples are FP32 (the wavefields), the phase tables and spectra are rounded to FP32, the products
// run on v_mfma_f32_32x32x2_f32 with FP32 accumulation over one slab, and every slab is added into
// FP64 accumulators (flush) or written as the FP32 source slice it becomes anyway (synth).
//
// MFMA 32x32x2 (one wavefront): lane l supplies A[i = l % 32][k = l / 32] and B[k = l / 32][j = l % 32];
// it receives D[i][j] for j = l % 32 and i = 8 * (r / 4) + 4 * (l / 32) + (r % 4), r = 0..15.
#include "engine.h"

namespace fd2d {

typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ f32x16 mfma32(float a, float b, f32x16 d)
{
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, d, 0, 0, 0);
}

__device__ __forceinline__ int d_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// ---------------------------------------------------------------------------------------------
// DFT flush. Workgroup = 64 nodes x 256 columns (128 frequencies); wavefront w owns columns
// 64w..64w+63 as two 32-row MFMA blocks against the two 32-node blocks.
// ---------------------------------------------------------------------------------------------
constexpr int kFX = 64;    // nodes per workgroup
constexpr int kFC = 256;   // table columns per workgroup
constexpr int kFU = 8;     // k-steps (sample pairs) per unrolled group

__global__ __launch_bounds__(256) void k_dft_mfma(double2* __restrict__ acc, const float* __restrict__ samp,
                                                  const float* __restrict__ etab, long P, int K, int C_pad, int cnt)
{
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const long x0 = (long)blockIdx.x * kFX;
    const int c0 = blockIdx.y * kFC + w * 64;
    const int half = lane >> 5, col = lane & 31;
    // clamped node columns of the two B blocks (loads stay inside the slab; stores are masked)
    const long xa = x0 + col < P ? x0 + col : P - 1;
    const long xb = x0 + 32 + col < P ? x0 + 32 + col : P - 1;
    const float* sa = samp + xa + (long)half * P;
    const float* sb = samp + xb + (long)half * P;
    const float* ea = etab + c0 + col + (long)half * C_pad;
    const float* eb = ea + 32;
    f32x16 d00 = {}, d01 = {}, d10 = {}, d11 = {};  // [column block][node block]
    const int pairs = cnt >> 1;
    int p = 0;
    for (; p + kFU <= pairs; p += kFU) {
        float a0[kFU], a1[kFU], b0[kFU], b1[kFU];
#pragma unroll
        for (int u = 0; u < kFU; ++u) {
            const long j = 2L * (p + u);
            b0[u] = sa[j * P];
            b1[u] = sb[j * P];
            a0[u] = ea[j * C_pad];
            a1[u] = eb[j * C_pad];
        }
#pragma unroll
        for (int u = 0; u < kFU; ++u) {
            d00 = mfma32(a0[u], b0[u], d00);
            d01 = mfma32(a0[u], b1[u], d01);
            d10 = mfma32(a1[u], b0[u], d10);
            d11 = mfma32(a1[u], b1[u], d11);
        }
    }
    for (; 2 * p < cnt; ++p) {  // remaining pairs; an odd last sample pairs with zero
        const long j = 2L * p;
        const bool ok = j + half < cnt;
        const float b0 = ok ? sa[j * P] : 0.f, b1 = ok ? sb[j * P] : 0.f;
        const float a0 = ok ? ea[j * C_pad] : 0.f, a1 = ok ? eb[j * C_pad] : 0.f;
        d00 = mfma32(a0, b0, d00);
        d01 = mfma32(a0, b1, d01);
        d10 = mfma32(a1, b0, d10);
        d11 = mfma32(a1, b1, d11);
    }
    // rows come in (real, imaginary) pairs: column c0 + 32cb + row, row even
    auto put = [&](const f32x16& d, int cb, int nb) {
        const long x = x0 + 32 * nb + col;
        if (x >= P) return;
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const int k = (c0 + 32 * cb + d_row(r, lane)) >> 1;
            if (k < K) {
                double2* q = acc + (long)k * P + x;
                double2 v = *q;
                v.x += (double)d[r];
                v.y += (double)d[r + 1];
                *q = v;
            }
        }
    };
    put(d00, 0, 0);
    put(d01, 0, 1);
    put(d10, 1, 0);
    put(d11, 1, 1);
}

void launch_dft_mfma(double2* acc, const float* samp, const float* etab, size_t P, int K, int C_pad, int cnt,
                     hipStream_t s)
{
    if (cnt <= 0 || K <= 0) return;
    const unsigned nxb = (unsigned)((P + kFX - 1) / kFX);
    const unsigned ncg = (unsigned)(C_pad / kFC);
    hipLaunchKernelGGL(k_dft_mfma, dim3(nxb, ncg), dim3(256), 0, s, acc, samp, etab, (long)P, K, C_pad, cnt);
}

// ---------------------------------------------------------------------------------------------
// Synthesis. Workgroup = 64 nodes x 256 time slices; wavefront w owns slices 64w..64w+63 as two
// 32-row MFMA blocks against the two 32-node blocks; the contraction runs over the C_pad spectrum
// rows (zero rows pad 2K up to a multiple of 256).
// ---------------------------------------------------------------------------------------------
constexpr int kSJ = 256;  // time slices per workgroup

__global__ __launch_bounds__(256) void k_synth_mfma(float* __restrict__ out, const float* __restrict__ a,
                                                    const float* __restrict__ et, long P, int C_pad, int J_pad,
                                                    int cnt)
{
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const long x0 = (long)blockIdx.x * kFX;
    const int j0 = blockIdx.y * kSJ + w * 64;
    const int half = lane >> 5, col = lane & 31;
    const long xa = x0 + col < P ? x0 + col : P - 1;
    const long xb = x0 + 32 + col < P ? x0 + 32 + col : P - 1;
    const float* pa = a + xa + (long)half * P;      // a[c][x], row c = 2q + half
    const float* pb = a + xb + (long)half * P;
    const float* ta = et + j0 + col + (long)half * J_pad;  // et[c][j]
    const float* tb = ta + 32;
    f32x16 d00 = {}, d01 = {}, d10 = {}, d11 = {};  // [slice block][node block]
    const int pairs = C_pad >> 1;  // C_pad is a multiple of 256
    for (int p = 0; p < pairs; p += kFU) {
        float a0[kFU], a1[kFU], b0[kFU], b1[kFU];
#pragma unroll
        for (int u = 0; u < kFU; ++u) {
            const long c = 2L * (p + u);
            b0[u] = pa[c * P];
            b1[u] = pb[c * P];
            a0[u] = ta[c * J_pad];
            a1[u] = tb[c * J_pad];
        }
#pragma unroll
        for (int u = 0; u < kFU; ++u) {
            d00 = mfma32(a0[u], b0[u], d00);
            d01 = mfma32(a0[u], b1[u], d01);
            d10 = mfma32(a1[u], b0[u], d10);
            d11 = mfma32(a1[u], b1[u], d11);
        }
    }
    auto put = [&](const f32x16& d, int jb, int nb) {
        const long x = x0 + 32 * nb + col;
        if (x >= P) return;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int j = j0 + 32 * jb + d_row(r, lane);
            if (j < cnt) out[(long)j * P + x] = d[r];
        }
    };
    put(d00, 0, 0);
    put(d01, 0, 1);
    put(d10, 1, 0);
    put(d11, 1, 1);
}

void launch_synth_mfma(float* out, const float* a, const float* et, size_t P, int C_pad, int J_pad, int cnt,
                       hipStream_t s)
{
    if (cnt <= 0) return;
    const unsigned nxb = (unsigned)((P + kFX - 1) / kFX);
    const unsigned njg = (unsigned)((cnt + kSJ - 1) / kSJ);
    hipLaunchKernelGGL(k_synth_mfma, dim3(nxb, njg), dim3(256), 0, s, out, a, et, (long)P, C_pad, J_pad, cnt);
}

// a[2k][x] = Re A[k][x], a[2k+1][x] = Im A[k][x] (FP32), zero rows 2K..C_pad-1
__global__ void k_spectrum_f32(float* __restrict__ a, const double2* __restrict__ A, long P, int K, int C_pad)
{
    const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (x >= P) return;
    if (k < K) {
        const double2 v = A[(long)k * P + x];
        a[(2L * k) * P + x] = (float)v.x;
        a[(2L * k + 1) * P + x] = (float)v.y;
    } else {
        a[(2L * k) * P + x] = 0.f;
        a[(2L * k + 1) * P + x] = 0.f;
    }
}

void launch_spectrum_f32(float* a, const double2* A, size_t P, int K, int C_pad, hipStream_t s)
{
    hipLaunchKernelGGL(k_spectrum_f32, dim3((unsigned)((P + 255) / 256), (unsigned)(C_pad / 2)), dim3(256), 0, s, a,
                       A, (long)P, K, C_pad);
}

}  // namespace fd2d
