// spectral.hip -- frequency-domain slab products of the engine on the matrix cores.
//
// Both contractions are tall-skinny products over the P grid nodes, with the complex phase
// factors of a slab of time samples split into real columns c = 2k (real part) and c = 2k + 1:
//   flush:  acc[k][x] += sum_j samp[j][x] * e[j][k]               (the DFT of a wavefield,
//           code/run_forward1_green.m:176-183, run_noise_adjoint.m:335-346)
//   synth:  out[j][x]  = sum_c et[c][j] * a[c][x]                  (the distributed source of the
//           sampled steps, run_forward2_correlation.m:195-216: Re(A e^{iwt}) = Ar*cr - Ai*ci)
// Samples and spectra are FP32, the phase factors are rounded to FP32. Every FP32 operand is split
// exactly into three bfloat16 parts, x = h + m + l (8 significant bits each, rounded to nearest),
// and a product is formed from the six part products down to 2^-16 of the leading one,
//   a b ~ ah bh + ah bm + am bh + am bm + ah bl + al bh,
// which leaves out terms below 2^-23 of |a b|, the size of an FP32 product's own rounding. The part
// products run on v_mfma_f32_32x32x16_bf16 (17x the rate of the FP32 v_mfma_f32_32x32x2_f32, so six
// of them take a third of its time) with FP32 accumulation over one slab, and every slab is added
// into FP64 accumulators with Neumaier compensation (flush) or written as the FP32 source slice it
// becomes anyway (synth). The
// phase tables are split once on the host; the samples and spectra as they are loaded.
//
// MFMA 32x32x16 (one wavefront): lane l supplies A[i = l % 32][k = 8 (l / 32) + v] and
// B[k = 8 (l / 32) + v][j = l % 32] for v = 0..7; it receives D[i][j] for j = l % 32 and
// i = 8 * (r / 4) + 4 * (l / 32) + (r % 4), r = 0..15.
#include "engine.h"

#include <algorithm>
#include <cstring>

namespace fd2d {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int d_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// the three bfloat16 parts of FP32 values: h = x rounded to 8 significant bits (to nearest, ties to
// even), m the same of x - h, l = x - h - m (exact: each difference is representable, and l has at
// most 8 bits left). Rounding to nearest keeps |m| <= 2^-8 |x| and |l| <= 2^-16 |x| with random
// signs, so the part products left out of a product (m l, l m, l l) stay below 2^-23 of it and
// carry no bias. split2 packs the parts of two values as the bfloat16 pairs of three 32-bit words.
struct Split8 {
    bf16x8 h, m, l;
};
__device__ __forceinline__ unsigned bf16_rne(unsigned u) { return (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u; }
__device__ __forceinline__ void split2(float v0, float v1, unsigned& h, unsigned& m, unsigned& l)
{
    unsigned hb[2], mb[2], lb[2];
    const float v[2] = {v0, v1};
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        hb[e] = bf16_rne(__float_as_uint(v[e]));
        const float r1 = v[e] - __uint_as_float(hb[e]);
        mb[e] = bf16_rne(__float_as_uint(r1));
        lb[e] = __float_as_uint(r1 - __uint_as_float(mb[e])) & 0xffff0000u;
    }
    h = (hb[0] >> 16) | hb[1];
    m = (mb[0] >> 16) | mb[1];
    l = (lb[0] >> 16) | lb[1];
}
// 8 consecutive bfloat16 values of a pre-split table (16-byte aligned)
__device__ __forceinline__ bf16x8 ld8(const unsigned short* p)
{
    return __builtin_bit_cast(bf16x8, *(const u32x4*)p);
}
__device__ __forceinline__ Split8 ld_split(const unsigned short* h, const unsigned short* m, const unsigned short* l,
                                           long o)
{
    return {ld8(h + o), ld8(m + o), ld8(l + o)};
}
__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 d)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, d, 0, 0, 0);
}
// d[p][nb] += A[p] B over the six part products, smallest first; the two column (or slice)
// blocks interleave
template <int NB>
__device__ __forceinline__ void mfma6(f32x16 (&d)[2][NB], const Split8 (&A)[2], const Split8& b, int nb)
{
#pragma unroll
    for (int t = 0; t < 6; ++t)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const Split8& a = A[p];
            switch (t) {
            case 0: d[p][nb] = mfma(a.l, b.h, d[p][nb]); break;
            case 1: d[p][nb] = mfma(a.h, b.l, d[p][nb]); break;
            case 2: d[p][nb] = mfma(a.m, b.m, d[p][nb]); break;
            case 3: d[p][nb] = mfma(a.m, b.h, d[p][nb]); break;
            case 4: d[p][nb] = mfma(a.h, b.m, d[p][nb]); break;
            default: d[p][nb] = mfma(a.h, b.h, d[p][nb]); break;
            }
        }
}

void split_bf16(float v, unsigned short& h, unsigned short& m, unsigned short& l)
{
    auto bits = [](float f) {
        unsigned u;
        std::memcpy(&u, &f, 4);
        return u;
    };
    auto value = [](unsigned u) {
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    };
    auto rne = [](unsigned u) { return (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u; };
    const unsigned hb = rne(bits(v));
    const float r1 = v - value(hb);
    const unsigned mb = rne(bits(r1));
    const unsigned lb = bits(r1 - value(mb)) & 0xffff0000u;
    h = (unsigned short)(hb >> 16);
    m = (unsigned short)(mb >> 16);
    l = (unsigned short)(lb >> 16);
}

// ---------------------------------------------------------------------------------------------
// Staging of the FP32 operand shared by a workgroup's wavefronts. Per chunk of 16 contraction
// rows, the rows of its 128 nodes are loaded kPF chunks ahead (each thread some rows of one
// node), split into their three parts and written to LDS in the B-operand order of the MFMA: for
// node block nb, part p, lane (half, col): the 8 values of rows 8 half .. 8 half + 7 of node
// 32 nb + col. Two LDS buffers alternate, so one barrier per chunk orders the writes after the
// previous reads. Each wavefront multiplies its two 32-row table blocks with all four node blocks,
// so every table value it loads serves 128 nodes.
// ---------------------------------------------------------------------------------------------
constexpr int kCH = 16;   // contraction rows per chunk (one MFMA k-step)

// staging area [buf][node block (NB)][part][half][col], 16 bytes each
template <int NB>
__device__ __forceinline__ const u32x4* b_slot(const u32x4* Bs, int buf, int nb, int half, int col)
{
    return Bs + ((buf * NB + nb) * 3) * 64 + half * 32 + col;  // parts 64 apart
}
// split RPT consecutive rows v0 .. v0 + RPT - 1 at lane position (half, col) of node block nb and
// write their parts (RPT = 2, 4 or 8)
template <int NB, int RPT>
__device__ __forceinline__ void stage_rows(u32x4* Bs, int buf, int nb, int half, int col, int v0, const float (&r)[RPT])
{
    unsigned ph[RPT / 2], pm[RPT / 2], pl[RPT / 2];
#pragma unroll
    for (int q = 0; q < RPT / 2; ++q) split2(r[2 * q], r[2 * q + 1], ph[q], pm[q], pl[q]);
    unsigned* base = (unsigned*)b_slot<NB>(Bs, buf, nb, half, col) + v0 / 2;
#pragma unroll
    for (int q = 0; q < RPT / 2; ++q) {
        base[q] = ph[q];
        base[q + 64 * 4] = pm[q];
        base[q + 128 * 4] = pl[q];
    }
}
template <int NB>
__device__ __forceinline__ Split8 read_b(const u32x4* Bs, int buf, int nb, int lane)
{
    const u32x4* q = b_slot<NB>(Bs, buf, nb, lane >> 5, lane & 31);
    return {__builtin_bit_cast(bf16x8, q[0]), __builtin_bit_cast(bf16x8, q[64]), __builtin_bit_cast(bf16x8, q[128])};
}

// ---------------------------------------------------------------------------------------------
// DFT flush. Workgroup = 128 nodes x 256 columns (128 frequencies); wavefront w owns columns
// 64w..64w+63 as two 32-row MFMA blocks against the four 32-node blocks. Thread t stages rows
// 8 (t / 128) .. + 7 of every chunk of samples for node t % 128. Table parts are [C_pad][Jt]
// (samples contiguous), offset to the slab's first sample (a multiple of 16), loaded one chunk
// ahead.
// ---------------------------------------------------------------------------------------------
constexpr int kFX = 128;   // nodes per workgroup
constexpr int kNB = kFX / 32;
constexpr int kFC = 256;   // table columns per workgroup
constexpr int kPFD = 2;    // chunks of samples loaded ahead

// s + c += x with Neumaier's compensation: c collects the rounding error of every addition into s
// (exactly, by the two-sum of the larger and the smaller operand), so s + c keeps the sum of every x
// added to within twice the FP64 rounding of the sum itself, whatever the number of additions.
__device__ __forceinline__ void neumaier(double& s, double& c, double x)
{
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

__global__ __launch_bounds__(256, 2) void k_dft_bf16(double2* __restrict__ acc, double2* __restrict__ comp,
                                                  const float* __restrict__ samp,
                                                  const unsigned short* __restrict__ eh,
                                                  const unsigned short* __restrict__ em,
                                                  const unsigned short* __restrict__ el, long P, int K, int Jt,
                                                  int cnt, int first, double2* __restrict__ out, double scale,
                                                  int conj)
{
    __shared__ u32x4 Bs[2 * kNB * 3 * 64];  // [buf][node block][part][half][col]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long x0 = (long)blockIdx.x * kFX;
    const int c0 = blockIdx.y * kFC + w * 64;
    const int half = lane >> 5, col = lane & 31;
    // staging: node x0 + (tid % 128) (clamped: loads stay inside the slab, stores are masked)
    const int sn = tid & (kFX - 1), sh = tid >> 7;
    const float* sp = samp + (x0 + sn < P ? x0 + sn : P - 1);
    const int nch = (cnt + kCH - 1) / kCH;
    // Loads are unconditional (rows past the slab read its last row, then count as zero samples
    // against the next slab's factors): a conditional load would make the compiler wait for every
    // load in flight before the first product of each chunk.
    auto fetch = [&](int q, float (&r)[8]) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int j = kCH * q + 8 * sh + i;
            const float v = sp[(long)(j < cnt ? j : cnt - 1) * P];
            r[i] = j < cnt ? v : 0.f;
        }
    };
    const long ra = (long)(c0 + col) * Jt + 8 * half, rb = ra + 32L * Jt;  // the lane's two table rows
    auto table = [&](int q, Split8 (&E)[2]) {
        E[0] = ld_split(eh, em, el, ra + kCH * q);
        E[1] = ld_split(eh, em, el, rb + kCH * q);
    };
    float raw[kPFD][8];
#pragma unroll
    for (int u = 0; u < kPFD; ++u) fetch(u < nch ? u : nch - 1, raw[u]);
    Split8 E[2];
    table(0, E);
    f32x16 d[2][kNB] = {};  // [column block][node block]
    for (int q0 = 0; q0 < nch; q0 += kPFD) {
#pragma unroll
        for (int u = 0; u < kPFD; ++u) {
            const int q = q0 + u;
            if (q >= nch) break;  // uniform
            stage_rows<kNB, 8>(Bs, q & 1, sn >> 5, sh, sn & 31, 0, raw[u]);
            fetch(q + kPFD < nch ? q + kPFD : nch - 1, raw[u]);
            __syncthreads();
            Split8 En[2];
            table(q + 1 < nch ? q + 1 : q, En);
            __builtin_amdgcn_sched_barrier(0);  // issue the next chunk's table loads before the products
#pragma unroll
            for (int nb = 0; nb < kNB; ++nb) mfma6<kNB>(d, E, read_b<kNB>(Bs, q & 1, nb, lane), nb);
            E[0] = En[0];
            E[1] = En[1];
        }
    }
    // rows come in (real, imaginary) pairs: column c0 + 32cb + row, row even. The slab sums are
    // exact in FP64 and go into the FP64 accumulators with Neumaier compensation (sum + its
    // compensation term, both read for one node block, clamped addresses, before any is written
    // back). The first slab of a series starts from zero sums without reading them; the last one
    // writes the finished value scale * (sum + compensation), conjugated if asked, to `out`
    // instead of the accumulators (the operations of a separate finishing pass, in its order).
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb) {
            double2 v[8], cv[8];
            long o[8];
            bool ok[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const long x = x0 + 32 * nb + col;
                const int k = (c0 + 32 * cb + d_row(2 * r, lane)) >> 1;
                ok[r] = x < P && k < K;
                o[r] = ok[r] ? (long)k * P + x : 0;
                v[r] = first ? make_double2(0.0, 0.0) : acc[o[r]];
                cv[r] = first ? make_double2(0.0, 0.0) : comp[o[r]];
            }
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (ok[r]) {
                    neumaier(v[r].x, cv[r].x, (double)d[cb][nb][2 * r]);
                    neumaier(v[r].y, cv[r].y, (double)d[cb][nb][2 * r + 1]);
                    if (out) {
                        const double re = v[r].x + cv[r].x, im = v[r].y + cv[r].y;
                        out[o[r]] = make_double2(scale * re, conj ? -(scale * im) : scale * im);
                    } else {
                        acc[o[r]] = v[r];
                        comp[o[r]] = cv[r];
                    }
                }
        }
}

void launch_dft_bf16(double2* acc, double2* comp, const float* samp, const unsigned short* eh, const unsigned short* em,
                     const unsigned short* el, size_t P, int K, int C_pad, int Jt, int cnt, int first, double2* out,
                     double scale, int conj, hipStream_t s)
{
    if (cnt <= 0 || K <= 0) return;
    const unsigned nxb = (unsigned)((P + kFX - 1) / kFX);
    const unsigned ncg = (unsigned)(C_pad / kFC);
    hipLaunchKernelGGL(k_dft_bf16, dim3(nxb, ncg), dim3(256), 0, s, acc, comp, samp, eh, em, el, (long)P, K, Jt,
                       cnt, first, out, scale, conj);
}

// ---------------------------------------------------------------------------------------------
// Synthesis. Workgroup = 128 nodes x 256 time slices (8 wavefronts); wavefront w owns slices
// 32w..32w+31 as one 32-row MFMA block against the four 32-node blocks, so each table value it
// loads serves 128 nodes with only 64 accumulator registers (wavefronts past cnt compute on slice 0
// and store nothing, so that every load stays unconditional). Thread t stages rows
// 4 (t / 128) .. + 3 of every chunk of spectrum rows for node t % 128;
// the contraction runs over the C_pad spectrum rows (zero rows pad 2K up to a multiple of 256).
// Table parts are [J_pad][C_pad] (spectrum rows contiguous) per slab, loaded at the start of each
// chunk.
// ---------------------------------------------------------------------------------------------
constexpr int kSX = 128;   // nodes per workgroup
constexpr int kSNB = kSX / 32;
constexpr int kSJ = 256;   // time slices per workgroup
constexpr int kSW = kSJ / 32;  // wavefronts per workgroup
constexpr int kPFS = 1;    // chunks of spectrum rows loaded ahead (two workgroups per CU hide the rest)

// d0 += A B0 and d1 += A B1 over the six part products, the two chains interleaved
__device__ __forceinline__ void mfma6x2(f32x16& d0, f32x16& d1, const Split8& a, const Split8& b0, const Split8& b1)
{
#pragma unroll
    for (int t = 0; t < 6; ++t)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const Split8& b = q ? b1 : b0;
            f32x16& d = q ? d1 : d0;
            switch (t) {
            case 0: d = mfma(a.l, b.h, d); break;
            case 1: d = mfma(a.h, b.l, d); break;
            case 2: d = mfma(a.m, b.m, d); break;
            case 3: d = mfma(a.m, b.h, d); break;
            case 4: d = mfma(a.h, b.m, d); break;
            default: d = mfma(a.h, b.h, d); break;
            }
        }
}

__global__ __launch_bounds__(512, 4) void k_synth_bf16(float* __restrict__ out, const float* __restrict__ a,
                                                     const unsigned short* __restrict__ th,
                                                     const unsigned short* __restrict__ tm,
                                                     const unsigned short* __restrict__ tl, long P, int C_pad,
                                                     int cnt, int nxb, int njg, int per_xcd)
{
    __shared__ u32x4 Bs[2 * kSNB * 3 * 64];  // [buf][node block][part][half][col]
    // workgroup g runs on XCD g % 8; there the njg slice groups of a node block follow each other,
    // so the block's spectrum rows come from HBM once and from that XCD's L2 for the other groups
    const int g = blockIdx.x, k = g / kXCD;
    const int xb = (g % kXCD) * per_xcd + k / njg, sg = k % njg;
    if (xb >= nxb) return;  // workgroup-uniform
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long x0 = (long)xb * kSX;
    const int j0 = sg * kSJ + w * 32;
    const bool active = j0 < cnt;  // wave-uniform
    const int half = lane >> 5, col = lane & 31;
    const int sn = tid & (kSX - 1), sr = tid >> 7;  // staging: node, group of 4 rows
    const float* sp = a + (x0 + sn < P ? x0 + sn : P - 1);  // a[c][x]
    const int nch = C_pad / kCH;
    auto fetch = [&](int q, float (&r)[4]) {
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] = sp[(long)(kCH * q + 4 * sr + i) * P];
    };
    // the lane's slice (slice 0 for the wavefronts past cnt)
    const long ta = (long)((active ? j0 : 0) + col) * C_pad + 8 * half;
    float raw[kPFS][4];
#pragma unroll
    for (int u = 0; u < kPFS; ++u) fetch(u < nch ? u : nch - 1, raw[u]);
    f32x16 d[kSNB] = {};  // [node block]
    for (int q0 = 0; q0 < nch; q0 += kPFS) {
#pragma unroll
        for (int u = 0; u < kPFS; ++u) {
            const int q = q0 + u;
            if (q >= nch) break;  // uniform
            const Split8 T = ld_split(th, tm, tl, ta + kCH * q);  // L2-resident; in flight across the staging
            __builtin_amdgcn_sched_barrier(0);
            stage_rows<kSNB, 4>(Bs, q & 1, sn >> 5, sr >> 1, sn & 31, 4 * (sr & 1), raw[u]);
            fetch(q + kPFS < nch ? q + kPFS : nch - 1, raw[u]);
            __syncthreads();
#pragma unroll
            for (int nb = 0; nb < kSNB; nb += 2)
                mfma6x2(d[nb], d[nb + 1], T, read_b<kSNB>(Bs, q & 1, nb, lane), read_b<kSNB>(Bs, q & 1, nb + 1, lane));
        }
    }
    if (!active) return;
#pragma unroll
    for (int nb = 0; nb < kSNB; ++nb) {
        const long x = x0 + 32 * nb + col;
        if (x >= P) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int j = j0 + d_row(r, lane);
            if (j < cnt) out[(long)j * P + x] = d[nb][r];
        }
    }
}

void launch_synth_bf16(float* out, const float* a, const unsigned short* th, const unsigned short* tm,
                       const unsigned short* tl, size_t P, int C_pad, int cnt, hipStream_t s)
{
    if (cnt <= 0) return;
    const int nxb = (int)((P + kSX - 1) / kSX);
    const int njg = (cnt + kSJ - 1) / kSJ;
    const int per_xcd = (nxb + kXCD - 1) / kXCD;
    hipLaunchKernelGGL(k_synth_bf16, dim3((unsigned)(kXCD * per_xcd * njg)), dim3(kSW * 64), 0, s, out, a, th, tm, tl,
                       (long)P, C_pad, cnt, nxb, njg, per_xcd);
}


}  // namespace fd2d
