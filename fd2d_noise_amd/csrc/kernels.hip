// kernels.hip -- gfx950 kernels of the 2-D SH staggered-grid engine.
//
// One time step is one launch of k_tile<R,MODE> over the nx x nz grid (64 x 32 nodes per
// workgroup, one wavefront per row slice, so every load and store of a row is 256 contiguous
// bytes). Each workgroup recomputes the velocity on a halo of its tile, so the step needs no second
// launch:
//   velocity: DS = div(sigma) (zero frames of code/propagation/div_s.m), point sources, synthesised
//             distributed source, v = (v + dt/rho*DS) * taper, u += v*dt, snapshot of the
//             pre-update field, DFT sample store, receiver gather;
//   stress:   sigma += dt*mu * D(v) (zero frames of dx_v.m / dz_v.m), Born terms.
// The structure-kernel product K_mu -= fw_nth * (D u . D u_fwd) runs as k_kmu after the steps that
// take it. The frequency-domain work is blocked over time: DFT samples and synthesised source slices
// are staged as FP32 slabs and contracted with FP64 phase tables in separate launches.
#include "engine.h"

namespace fd2d {

// ---------------------------------------------------------------------------------------------
// Fused time step
// ---------------------------------------------------------------------------------------------
// One workgroup (4 wavefronts) advances one kTX x kTileZ tile by one step:
//   phase 1 issues, in one burst, the loads of sigma over the tile plus the halo that the velocity
//           stencil of the halo needs, the velocity operands and (structure-kernel steps) the
//           forward snapshot, then stores them to LDS;
//   phase 2 computes the velocity (and displacement) on the tile plus an (R-1, R) halo in both
//           directions -- the halo is recomputed, bit-identically, by the neighbouring tiles --
//           and writes the outputs of the tile's own nodes;
//   phase 3 updates sigma (and K_mu) on the tile.
// Lane l of every wavefront owns column i0 + l; wavefront w owns rows w, w+4, ... of each region,
// so row offsets are wave-uniform and LDS row offsets fold into instruction immediates. The few
// halo columns (2R-1 of the velocity region, 2(2R-1) of the sigma_xy region) are handled as one
// packed list. Global loads are raw buffer loads with the whole byte offset in the VGPR, so the
// hardware range check covers every load (out-of-range reads return 0).
//
// Tiles are dealt to the XCDs in contiguous bands (tile = (g % 8) * tiles_per_xcd + g / 8), so a
// tile's x-neighbour, and a tile row later its z-neighbour, read the shared halo lines from the
// same L2.
//
// MODE 0: Green run (no displacement); 1: with displacement (correlation and adjoint runs).
__device__ __forceinline__ int node_key(int x, int z) { return ((z + 8) << 16) | (x + 8); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, int off)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

template <int R, int MODE>
__global__ __launch_bounds__(256) void k_tile(StepArgs a)
{
    constexpr bool HAS_U = MODE >= 1;
    constexpr int TZ = kTileZ;
    constexpr int WV = kTX + 2 * R - 1;     // velocity region: columns i0-(R-1) .. i0+63+R
    constexpr int HV = TZ + 2 * R - 1;      //                  rows    j0-(R-1) .. j0+TZ-1+R
    constexpr int WSX = WV + 2 * R - 1;     // sigma_xy region: columns i0-(2R-1) .. i0+63+2R-1
    constexpr int HSZ = HV + 2 * R - 1;     // sigma_zy region: rows    j0-(2R-1) .. j0+TZ-1+2R-1
    constexpr int HCV = 2 * R - 1;          // halo columns of the velocity / sigma_zy regions
    constexpr int HCS = 2 * (2 * R - 1);    // halo columns of the sigma_xy region
    __shared__ float sx[HV * WSX];
    __shared__ float sz[HSZ * WV];
    __shared__ float vv[HV * WV];

    const int g = blockIdx.x;
    const int t = (g % kXCD) * a.tiles_per_xcd + g / kXCD;
    if (t >= a.n_tiles) return;
    const int i0 = (t % a.tiles_x) * kTX, j0 = (t / a.tiles_x) * TZ;
    const int nx = a.nx, nz = a.nz, pitch = a.pitch;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x = i0 + lane;                 // this lane's own column
    const bool xg = x < nx;

    const unsigned pb = a.padded_bytes;
    const auto r_sxy = rsrc(a.sxy_in - a.origin, pb), r_szy = rsrc(a.szy_in - a.origin, pb);
    const auto r_v = rsrc(a.v_in - a.origin, pb), r_bv = rsrc(a.bv - a.origin, pb);
    const auto r_u = rsrc(HAS_U ? a.u_in - a.origin : a.v_in - a.origin, pb);
    const auto r_dm = rsrc(a.dtmu - a.origin, pb);
    // padded byte offset of node (xx, zz); zz >= -kGhost and xx >= -kGhostX on every use
    auto po = [&](int xx, int zz) { return ((zz + kGhost) * pitch + xx + kGhostX) * 4; };
    auto co = [&](int xx, int zz) { return (zz * nx + xx) * 4; };  // compact byte offset (in grid)
    auto ingrid = [&](int xx, int zz) { return xx >= 0 && xx < nx && zz >= 0 && zz < nz; };
    const float* syn = a.syn;

    // ---------------- phase 1
    constexpr int NRV = (HV + 3) / 4, NRZ = (HSZ + 3) / 4;
    constexpr int NHS = (HV * HCS + 255) / 256, NHZ = (HSZ * HCV + 255) / 256, NHV = (HV * HCV + 255) / 256;
    float msx[NRV], msz[NRZ];
    float mv[NRV], mb[NRV], mu[NRV], msy[NRV];
    float hsx[NHS], hsz[NHZ];
    float hv[NHV], hb[NHV], hu[NHV], hsy[NHV], ht[NHV];
    const float axl = xg ? a.ax[x] : 0.f;
#pragma unroll
    for (int k = 0; k < NRV; ++k) {
        const int r = w + 4 * k, z = j0 - (R - 1) + r;
        const bool live = r < HV;
        const int o = po(x, z);
        msx[k] = live ? bload(r_sxy, o) : 0.f;
        const bool in = live && xg && z >= 0 && z < nz;
        mv[k] = in ? bload(r_v, o) : 0.f;
        mb[k] = in ? bload(r_bv, o) : 0.f;
        mu[k] = (HAS_U && in) ? bload(r_u, o) : 0.f;
        msy[k] = (in && syn) ? syn[co(x, z) / 4] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NRZ; ++k) {
        const int r = w + 4 * k, z = j0 - (2 * R - 1) + r;
        msz[k] = r < HSZ ? bload(r_szy, po(x, z)) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NHS; ++k) {
        const int idx = tid + 256 * k;
        const int r = idx / HCS, h = idx - r * HCS;
        const int c = h < 2 * R - 1 ? h : h + kTX;            // sigma_xy region column
        const int xx = i0 - (2 * R - 1) + c, zz = j0 - (R - 1) + r;
        hsx[k] = idx < HV * HCS ? bload(r_sxy, po(xx, zz)) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NHZ; ++k) {
        const int idx = tid + 256 * k;
        const int r = idx / HCV, h = idx - r * HCV;
        const int c = h < R - 1 ? h : h + kTX;                // velocity-region column
        const int xx = i0 - (R - 1) + c, zz = j0 - (2 * R - 1) + r;
        hsz[k] = idx < HSZ * HCV ? bload(r_szy, po(xx, zz)) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NHV; ++k) {
        const int idx = tid + 256 * k;
        const int r = idx / HCV, h = idx - r * HCV;
        const int c = h < R - 1 ? h : h + kTX;
        const int xx = i0 - (R - 1) + c, zz = j0 - (R - 1) + r;
        const bool in = idx < HV * HCV && ingrid(xx, zz);
        const int o = po(xx, zz);
        hv[k] = in ? bload(r_v, o) : 0.f;
        hb[k] = in ? bload(r_bv, o) : 0.f;
        hu[k] = (HAS_U && in) ? bload(r_u, o) : 0.f;
        hsy[k] = (in && syn) ? syn[co(xx, zz) / 4] : 0.f;
        ht[k] = in ? a.ax[xx] * a.az[zz] : 0.f;
    }
    // LDS stores
#pragma unroll
    for (int k = 0; k < NRV; ++k) {
        const int r = w + 4 * k;
        if (r < HV) sx[r * WSX + lane + 2 * R - 1] = msx[k];
    }
#pragma unroll
    for (int k = 0; k < NRZ; ++k) {
        const int r = w + 4 * k;
        if (r < HSZ) sz[r * WV + lane + R - 1] = msz[k];
    }
#pragma unroll
    for (int k = 0; k < NHS; ++k) {
        const int idx = tid + 256 * k;
        const int r = idx / HCS, h = idx - r * HCS;
        if (idx < HV * HCS) sx[r * WSX + (h < 2 * R - 1 ? h : h + kTX)] = hsx[k];
    }
#pragma unroll
    for (int k = 0; k < NHZ; ++k) {
        const int idx = tid + 256 * k;
        const int r = idx / HCV, h = idx - r * HCV;
        if (idx < HSZ * HCV) sz[r * WV + (h < R - 1 ? h : h + kTX)] = hsz[k];
    }
    __syncthreads();

    // ---------------- phase 2
    const int e0 = a.ps_off ? a.ps_off[t] : 0, e1 = a.ps_off ? a.ps_off[t + 1] : 0;
    const int f0 = a.rec_off ? a.rec_off[t] : 0, f1 = a.rec_off ? a.rec_off[t + 1] : 0;
    // velocity at (xx, zz) = region cell (r, c) from the staged sigma; the point sources of the node
    auto divs = [&](int xx, int zz, int r, int c) {
        float ds = 0.f;
        if (xx >= R && xx <= nx - R - 1) {  // div_s x-part frame
            const float* row = sx + r * WSX + c + R;
            float s = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m) s += a.cx[m - 1] * (row[m - 1] - row[-m]);
            ds = s;
        }
        if (zz >= R && zz <= nz - R - 1) {  // div_s z-part frame
            const float* col = sz + (r + R) * WV + c;
            float s = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m) s += a.cz[m - 1] * (col[(m - 1) * WV] - col[-m * WV]);
            ds += s;
        }
        if (e1 > e0) {
            const int key = node_key(xx, zz);
            for (int e = e0; e < e1; ++e) {
                const int2 en = a.ps_ent[e];
                if (en.x == key) ds += a.ps_amp ? a.ps_amp[en.y] : a.ps_const;
            }
        }
        return ds;
    };
#pragma unroll
    for (int k = 0; k < NRV; ++k) {
        const int r = w + 4 * k;
        if (r >= HV) break;
        const int z = j0 - (R - 1) + r, c = lane + R - 1;
        float vn = 0.f;
        if (xg && z >= 0 && z < nz) {
            const float ds = divs(x, z, r, c) + msy[k];
            vn = (mv[k] + mb[k] * ds) * (axl * a.az[z]);
            const float un = HAS_U ? mu[k] + vn * a.dt : 0.f;
            if (r >= R - 1 && r < R - 1 + TZ) {  // own node: outputs
                const long p = (long)z * pitch + x;
                const long cc = (long)z * nx + x;
                a.v_out[p] = vn;
                float out = vn;
                if (HAS_U) {
                    a.u_out[p] = un;
                    if (a.snap) a.snap[cc] = mu[k];
                    out = un;
                } else if (a.snap) {
                    a.snap[cc] = mv[k];
                }
                if (a.dsamp) a.dsamp[cc] = out;
                if (f1 > f0) {
                    const int key = node_key(x, z);
                    for (int e = f0; e < f1; ++e) {
                        const int2 en = a.rec_ent[e];
                        if (en.x == key) a.seis_col[en.y] = (double)out;
                    }
                }
            }
        }
        vv[r * WV + c] = vn;
    }
#pragma unroll
    for (int k = 0; k < NHV; ++k) {
        const int idx = tid + 256 * k;
        if (idx >= HV * HCV) break;
        const int r = idx / HCV, h = idx - r * HCV;
        const int c = h < R - 1 ? h : h + kTX;
        const int xx = i0 - (R - 1) + c, zz = j0 - (R - 1) + r;
        float vn = 0.f;
        if (ingrid(xx, zz)) {
            const float ds = divs(xx, zz, r, c) + hsy[k];
            vn = (hv[k] + hb[k] * ds) * ht[k];
        }
        vv[r * WV + c] = vn;
    }
    __syncthreads();

    // ---------------- phase 3
    constexpr int NT = TZ / 4;
    float rm[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const int z = j0 + w + 4 * k;
        rm[k] = (xg && z < nz) ? bload(r_dm, po(x, z)) : 0.f;
    }
    const bool xin = x >= R - 1 && x <= nx - R - 1;  // dx_v frame
    const int c = lane + R - 1;
    const long nxl = nx;
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const int lz = w + 4 * k, z = j0 + lz;
        if (!xg || z >= nz) continue;
        const int r = lz + R - 1;
        const long p = (long)z * pitch + x;
        const long cc = (long)z * nx + x;
        const bool zin = z >= R - 1 && z <= nz - R - 1;  // dz_v frame
        const float dm = rm[k];
        const float* vrow = vv + r * WV + c;
        if (xin) {
            float e = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m) e += a.cx[m - 1] * (vrow[m] - vrow[1 - m]);
            float sn = sx[r * WSX + c + R] + dm * e;
            if (a.born_dmu) {
                float s = 0.f;
#pragma unroll
                for (int m = 1; m <= R; ++m)
                    s += a.cx[m - 1] * ((a.born_in0[cc + m] - a.born_in1[cc + m]) -
                                        (a.born_in0[cc - m + 1] - a.born_in1[cc - m + 1]));
                sn -= a.born_dmu[cc] * s;
            }
            a.sxy_out[p] = sn;
        }
        if (zin) {
            float e = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m) e += a.cz[m - 1] * (vrow[m * WV] - vrow[(1 - m) * WV]);
            float sn = sz[(r + R) * WV + c] + dm * e;
            if (a.born_dmu) {
                float s = 0.f;
#pragma unroll
                for (int m = 1; m <= R; ++m)
                    s += a.cz[m - 1] * ((a.born_in0[cc + m * nxl] - a.born_in1[cc + m * nxl]) -
                                        (a.born_in0[cc - (m - 1) * nxl] - a.born_in1[cc - (m - 1) * nxl]));
                sn -= a.born_dmu[cc] * s;
            }
            a.szy_out[p] = sn;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Two time steps per launch (temporal blocking)
// ---------------------------------------------------------------------------------------------
// A workgroup advances its kTX x kTileZ tile by steps n and n+1. It stages sigma of the state before
// step n on the tile plus a (4R-2)-node halo, computes the step-n velocity on the tile plus a
// (3R-2, 3R-1) halo (region V1), updates sigma in place where step n+1 needs it, computes the
// step-(n+1) velocity on the tile plus the (R-1, R) halo (V2) and finally sigma on the tile. The
// halo values are recomputed by the neighbouring tiles with the same operations, so a pair equals
// two single steps bit for bit; the compulsory HBM traffic per node is that of one step.
// Region coordinates relative to the tile origin (i0, j0), L1 = 3R-2, R1 = 3R-1, LS = 4R-2:
//   VV (V1): columns -L1 .. 63+R1, rows -L1 .. TZ-1+R1
//   SX:      columns -LS .. 63+LS, rows as VV
//   SZ:      columns as VV,        rows -LS .. TZ-1+LS
// Per-thread layout as in k_tile: lane l owns column i0 + l of every region's main block, wave w
// the rows w, w+4, ...; halo columns are packed lists over the 256 threads.
template <int R, int MODE>
__global__ __launch_bounds__(256) void k_pair(PairArgs pa)
{
    constexpr bool HAS_U = MODE >= 1;
    constexpr int TZ = kTileZ;
    constexpr int L1 = 3 * R - 2, R1 = 3 * R - 1, LS = 4 * R - 2;
    constexpr int W1 = kTX + L1 + R1, H1 = TZ + L1 + R1;
    constexpr int WSX = kTX + 2 * LS, HSZ = TZ + 2 * LS;
    constexpr int HC1 = L1 + R1;      // halo columns of VV and SZ
    constexpr int HCS = 2 * LS;       // halo columns of SX
    constexpr int V2R0 = 2 * R - 1, V2R1 = TZ + 4 * R - 3;    // VV rows of V2 (inclusive)
    constexpr int V2C0 = 2 * R - 1, V2C1 = kTX + 4 * R - 3;   // VV columns of V2 (inclusive)
    constexpr int HS1X = TZ + 2 * R - 1;     // rows of the step-n sigma_xy update (the V2 rows)
    constexpr int HS1Z = TZ + 4 * R - 2;     // rows of the step-n sigma_zy update: j0-(2R-1) ..
    constexpr int HCX1 = 2 * (2 * R - 1);    // halo columns of the sigma_xy update
    constexpr int HCZ1 = 2 * R - 1;          // halo columns of the sigma_zy update (V2 halo)
    __shared__ float SX[H1 * WSX];
    __shared__ float SZ[HSZ * W1];
    __shared__ float VV[H1 * W1];

    const StepArgs& a = pa.a;
    const int g = blockIdx.x;
    const int t = (g % kXCD) * a.tiles_per_xcd + g / kXCD;
    if (t >= a.n_tiles) return;
    const int i0 = (t % a.tiles_x) * kTX, j0 = (t / a.tiles_x) * TZ;
    const int nx = a.nx, nz = a.nz, pitch = a.pitch;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x = i0 + lane;
    const bool xg = x < nx;

    const unsigned pb = a.padded_bytes;
    const auto r_sxy = rsrc(a.sxy_in - a.origin, pb), r_szy = rsrc(a.szy_in - a.origin, pb);
    const auto r_v = rsrc(a.v_in - a.origin, pb), r_bv = rsrc(a.bv - a.origin, pb);
    const auto r_u = rsrc(HAS_U ? a.u_in - a.origin : a.v_in - a.origin, pb);
    const auto r_dm = rsrc(a.dtmu - a.origin, pb);
    // padded byte offset of node (xx, zz); zz >= -kGhost and xx >= -kGhostX on every use
    auto po = [&](int xx, int zz) { return ((zz + kGhost) * pitch + xx + kGhostX) * 4; };
    auto ingrid = [&](int xx, int zz) { return xx >= 0 && xx < nx && zz >= 0 && zz < nz; };
    // the synthesised source of the pair belongs to step n (a.syn) or to step n+1 (syn2)
    const float* syn = a.syn ? a.syn : pa.syn2;
    const bool syn1 = a.syn != nullptr;

    // ---------------- phase A: stage sigma, load the step-n velocity operands of V1
    constexpr int NM1 = (H1 + 3) / 4, NH1 = (H1 * HC1 + 255) / 256;
    constexpr int NMZ = (HSZ + 3) / 4;
    constexpr int NHX = (H1 * HCS + 255) / 256, NHZ = (HSZ * HC1 + 255) / 256;
    float mv[NM1], mb[NM1], msy[NM1], mu[NM1];
    float hv[NH1], hb[NH1], ht[NH1], hsy[NH1];
    {
        float sxm[NM1], sxh[NHX], szm[NMZ], szh[NHZ];
#pragma unroll
        for (int k = 0; k < NM1; ++k) {
            const int r = w + 4 * k;
            sxm[k] = r < H1 ? bload(r_sxy, po(x, j0 - L1 + r)) : 0.f;
        }
#pragma unroll
        for (int k = 0; k < NHX; ++k) {
            const int idx = tid + 256 * k;
            const int r = idx / HCS, h = idx - r * HCS;
            const int c = h < LS ? h : h + kTX;
            sxh[k] = idx < H1 * HCS ? bload(r_sxy, po(i0 - LS + c, j0 - L1 + r)) : 0.f;
        }
#pragma unroll
        for (int k = 0; k < NMZ; ++k) {
            const int r = w + 4 * k;
            szm[k] = r < HSZ ? bload(r_szy, po(x, j0 - LS + r)) : 0.f;
        }
#pragma unroll
        for (int k = 0; k < NHZ; ++k) {
            const int idx = tid + 256 * k;
            const int r = idx / HC1, h = idx - r * HC1;
            const int c = h < L1 ? h : h + kTX;
            szh[k] = idx < HSZ * HC1 ? bload(r_szy, po(i0 - L1 + c, j0 - LS + r)) : 0.f;
        }
#pragma unroll
        for (int k = 0; k < NM1; ++k) {
            const int r = w + 4 * k, z = j0 - L1 + r;
            const bool in = r < H1 && xg && z >= 0 && z < nz;
            const int o = po(x, z);
            mv[k] = in ? bload(r_v, o) : 0.f;
            mb[k] = in ? bload(r_bv, o) : 0.f;
            mu[k] = (HAS_U && in && r >= L1 && r < L1 + TZ) ? bload(r_u, o) : 0.f;
            msy[k] = (in && syn) ? syn[z * nx + x] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < NH1; ++k) {
            const int idx = tid + 256 * k;
            const int r = idx / HC1, h = idx - r * HC1;
            const int c = h < L1 ? h : h + kTX;
            const int xx = i0 - L1 + c, zz = j0 - L1 + r;
            const bool in = idx < H1 * HC1 && ingrid(xx, zz);
            const int o = po(xx, zz);
            hv[k] = in ? bload(r_v, o) : 0.f;
            hb[k] = in ? bload(r_bv, o) : 0.f;
            hsy[k] = (in && syn) ? syn[zz * nx + xx] : 0.f;
            ht[k] = in ? a.ax[xx] * a.az[zz] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < NM1; ++k) {
            const int r = w + 4 * k;
            if (r < H1) SX[r * WSX + lane + LS] = sxm[k];
        }
#pragma unroll
        for (int k = 0; k < NHX; ++k) {
            const int idx = tid + 256 * k;
            const int r = idx / HCS, h = idx - r * HCS;
            if (idx < H1 * HCS) SX[r * WSX + (h < LS ? h : h + kTX)] = sxh[k];
        }
#pragma unroll
        for (int k = 0; k < NMZ; ++k) {
            const int r = w + 4 * k;
            if (r < HSZ) SZ[r * W1 + lane + L1] = szm[k];
        }
#pragma unroll
        for (int k = 0; k < NHZ; ++k) {
            const int idx = tid + 256 * k;
            const int r = idx / HC1, h = idx - r * HC1;
            if (idx < HSZ * HC1) SZ[r * W1 + (h < L1 ? h : h + kTX)] = szh[k];
        }
    }
    __syncthreads();

    // ---------------- phase B: step-n velocity on V1
    const int e0 = a.ps_off ? a.ps_off[t] : 0, e1 = a.ps_off ? a.ps_off[t + 1] : 0;
    const int f0 = a.rec_off ? a.rec_off[t] : 0, f1 = a.rec_off ? a.rec_off[t + 1] : 0;
    auto divs = [&](int xx, int zz, int r, int c, const float* amp, float cst) {
        float ds = 0.f;
        if (xx >= R && xx <= nx - R - 1) {  // div_s x-part frame
            const float* row = SX + r * WSX + c + R;
            float s = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m) s += a.cx[m - 1] * (row[m - 1] - row[-m]);
            ds = s;
        }
        if (zz >= R && zz <= nz - R - 1) {  // div_s z-part frame
            const float* col = SZ + (r + R) * W1 + c;
            float s = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m) s += a.cz[m - 1] * (col[(m - 1) * W1] - col[-m * W1]);
            ds += s;
        }
        if (e1 > e0) {
            const int key = node_key(xx, zz);
            for (int e = e0; e < e1; ++e) {
                const int2 en = a.ps_ent[e];
                if (en.x == key) ds += amp ? amp[en.y] : cst;
            }
        }
        return ds;
    };
    auto gather = [&](int xx, int zz, float val, double* col) {
        if (f1 > f0 && col) {
            const int key = node_key(xx, zz);
            for (int e = f0; e < f1; ++e) {
                const int2 en = a.rec_ent[e];
                if (en.x == key) col[en.y] = (double)val;
            }
        }
    };
    // dt*mu of the sigma updates, prefetched: main columns over the sigma_zy rows (which contain
    // the sigma_xy rows and the tile), and the two packed halo lists
    constexpr int NMD = (HS1Z + 3) / 4;
    constexpr int NDX = (HS1X * HCX1 + 255) / 256, NDZ = (HS1Z * HCZ1 + 255) / 256;
    float dmm[NMD], dmx[NDX], dmz[NDZ];
#pragma unroll
    for (int k = 0; k < NMD; ++k) {
        const int q = w + 4 * k, z = j0 - (2 * R - 1) + q;
        dmm[k] = (q < HS1Z && xg && z >= 0 && z < nz) ? bload(r_dm, po(x, z)) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NDX; ++k) {
        const int idx = tid + 256 * k;
        const int r = idx / HCX1, h = idx - r * HCX1;
        const int xx = h < 2 * R - 1 ? i0 - (2 * R - 1) + h : i0 + kTX + (h - (2 * R - 1));
        const int zz = j0 - (R - 1) + r;
        dmx[k] = (idx < HS1X * HCX1 && ingrid(xx, zz)) ? bload(r_dm, po(xx, zz)) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NDZ; ++k) {
        const int idx = tid + 256 * k;
        const int r = idx / HCZ1, h = idx - r * HCZ1;
        const int xx = h < R - 1 ? i0 - (R - 1) + h : i0 + kTX + (h - (R - 1));
        const int zz = j0 - (2 * R - 1) + r;
        dmz[k] = (idx < HS1Z * HCZ1 && ingrid(xx, zz)) ? bload(r_dm, po(xx, zz)) : 0.f;
    }

    const float axl = xg ? a.ax[x] : 0.f;
    float un1[NM1];  // displacement after step n at the tile's own rows
#pragma unroll
    for (int k = 0; k < NM1; ++k) {
        const int r = w + 4 * k;
        un1[k] = 0.f;
        if (r >= H1) break;
        const int z = j0 - L1 + r, c = lane + L1;
        float vn = 0.f;
        if (xg && z >= 0 && z < nz) {
            const float ds = divs(x, z, r, c, a.ps_amp, a.ps_const) + (syn1 ? msy[k] : 0.f);
            vn = (mv[k] + mb[k] * ds) * (axl * a.az[z]);
            if (r >= L1 && r < L1 + TZ) {  // own node: outputs of step n
                const float un = HAS_U ? mu[k] + vn * a.dt : 0.f;
                un1[k] = un;
                const long cc = (long)z * nx + x;
                float out = vn;
                if (HAS_U) {
                    if (a.snap) a.snap[cc] = mu[k];
                    out = un;
                } else if (a.snap) {
                    a.snap[cc] = mv[k];
                }
                if (a.dsamp) a.dsamp[cc] = out;
                gather(x, z, out, a.seis_col);
            }
        }
        VV[r * W1 + c] = vn;
    }
#pragma unroll
    for (int k = 0; k < NH1; ++k) {
        const int idx = tid + 256 * k;
        if (idx >= H1 * HC1) break;
        const int r = idx / HC1, h = idx - r * HC1;
        const int c = h < L1 ? h : h + kTX;
        const int xx = i0 - L1 + c, zz = j0 - L1 + r;
        float vn = 0.f;
        if (ingrid(xx, zz)) {
            const float ds = divs(xx, zz, r, c, a.ps_amp, a.ps_const) + (syn1 ? hsy[k] : 0.f);
            vn = (hv[k] + hb[k] * ds) * ht[k];
        }
        VV[r * W1 + c] = vn;
    }
    __syncthreads();

    // ---------------- phase C: step-n sigma, in place, where step n+1 needs it
    const bool xin = x >= R - 1 && x <= nx - R - 1;  // dx_v frame
    auto dxv = [&](int rv, int c) {
        const float* vrow = VV + rv * W1 + c;
        float e = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) e += a.cx[m - 1] * (vrow[m] - vrow[1 - m]);
        return e;
    };
    auto dzv = [&](int rv, int c) {
        const float* vcol = VV + rv * W1 + c;
        float e = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) e += a.cz[m - 1] * (vcol[m * W1] - vcol[(1 - m) * W1]);
        return e;
    };
#pragma unroll
    for (int k = 0; k < NMD; ++k) {
        const int q = w + 4 * k;
        if (q >= HS1Z) break;
        const int z = j0 - (2 * R - 1) + q;
        if (!xg || z < 0 || z >= nz) continue;
        const int rv = z - (j0 - L1), c = lane + L1;
        if (z >= R - 1 && z <= nz - R - 1) SZ[(rv + R) * W1 + c] += dmm[k] * dzv(rv, c);  // dz_v frame
        if (rv >= V2R0 && rv <= V2R1 && xin) SX[rv * WSX + c + R] += dmm[k] * dxv(rv, c);
    }
#pragma unroll
    for (int k = 0; k < NDX; ++k) {
        const int idx = tid + 256 * k;
        if (idx >= HS1X * HCX1) break;
        const int r = idx / HCX1, h = idx - r * HCX1;
        const int xx = h < 2 * R - 1 ? i0 - (2 * R - 1) + h : i0 + kTX + (h - (2 * R - 1));
        const int zz = j0 - (R - 1) + r;
        if (ingrid(xx, zz) && xx >= R - 1 && xx <= nx - R - 1) {
            const int rv = zz - (j0 - L1), cv = xx - (i0 - L1);
            SX[rv * WSX + cv + R] += dmx[k] * dxv(rv, cv);
        }
    }
#pragma unroll
    for (int k = 0; k < NDZ; ++k) {
        const int idx = tid + 256 * k;
        if (idx >= HS1Z * HCZ1) break;
        const int r = idx / HCZ1, h = idx - r * HCZ1;
        const int xx = h < R - 1 ? i0 - (R - 1) + h : i0 + kTX + (h - (R - 1));
        const int zz = j0 - (2 * R - 1) + r;
        if (ingrid(xx, zz) && zz >= R - 1 && zz <= nz - R - 1) {
            const int rv = zz - (j0 - L1), cv = xx - (i0 - L1);
            SZ[(rv + R) * W1 + cv] += dmz[k] * dzv(rv, cv);
        }
    }
    __syncthreads();

    // ---------------- phase D: step-(n+1) velocity on V2, in place
#pragma unroll
    for (int k = 0; k < NM1; ++k) {
        const int r = w + 4 * k;
        if (r >= H1) break;
        if (r < V2R0 || r > V2R1) continue;
        const int z = j0 - L1 + r, c = lane + L1;
        if (!xg || z < 0 || z >= nz) continue;  // stays 0
        const float v1 = VV[r * W1 + c];
        const float ds = divs(x, z, r, c, pa.ps_amp2, pa.ps_const2) + (syn1 ? 0.f : msy[k]);
        const float vn = (v1 + mb[k] * ds) * (axl * a.az[z]);
        if (r >= L1 && r < L1 + TZ) {  // own node: outputs of step n+1 and the new state
            const long p = (long)z * pitch + x;
            const long cc = (long)z * nx + x;
            a.v_out[p] = vn;
            float out = vn;
            if (HAS_U) {
                const float un = un1[k] + vn * a.dt;
                a.u_out[p] = un;
                if (pa.snap2) pa.snap2[cc] = un1[k];
                out = un;
            } else if (pa.snap2) {
                pa.snap2[cc] = v1;
            }
            if (pa.dsamp2) pa.dsamp2[cc] = out;
            gather(x, z, out, pa.seis_col2);
        }
        VV[r * W1 + c] = vn;
    }
#pragma unroll
    for (int k = 0; k < NH1; ++k) {
        const int idx = tid + 256 * k;
        if (idx >= H1 * HC1) break;
        const int r = idx / HC1, h = idx - r * HC1;
        const int c = h < L1 ? h : h + kTX;
        if (r < V2R0 || r > V2R1 || c < V2C0 || c > V2C1) continue;
        const int xx = i0 - L1 + c, zz = j0 - L1 + r;
        if (!ingrid(xx, zz)) continue;
        const float ds = divs(xx, zz, r, c, pa.ps_amp2, pa.ps_const2) + (syn1 ? 0.f : hsy[k]);
        VV[r * W1 + c] = (VV[r * W1 + c] + hb[k] * ds) * ht[k];
    }
    __syncthreads();

    // ---------------- phase E: step-(n+1) sigma on the tile
#pragma unroll
    for (int k = 0; k < NMD; ++k) {
        const int q = w + 4 * k;
        if (q >= HS1Z) break;
        const int z = j0 - (2 * R - 1) + q;
        if (z < j0 || z >= j0 + TZ || !xg || z >= nz) continue;
        const int rv = z - (j0 - L1), c = lane + L1;
        const long p = (long)z * pitch + x;
        if (xin) a.sxy_out[p] = SX[rv * WSX + c + R] + dmm[k] * dxv(rv, c);
        if (z >= R - 1 && z <= nz - R - 1) a.szy_out[p] = SZ[(rv + R) * W1 + c] + dmm[k] * dzv(rv, c);
    }
}

// Structure-kernel product of run_noise_adjoint.m:324-331 after a step that stores it:
// K_mu -= fw_nth * (dx_v(u) .* dx_v(u_fwd) + dz_v(u) .* dz_v(u_fwd)) with u the new displacement
// (padded) and u_fwd the forward snapshot (compact); the products are formed in FP64.
template <int R>
__global__ __launch_bounds__(256) void k_kmu(StepArgs a, const float* __restrict__ u, const float* __restrict__ fwd,
                                             double* __restrict__ kmu, double scale)
{
    const int x = blockIdx.x * kTX + threadIdx.x, z = blockIdx.y * 4 + threadIdx.y;
    if (x >= a.nx || z >= a.nz) return;
    const bool xin = x >= R - 1 && x <= a.nx - R - 1, zin = z >= R - 1 && z <= a.nz - R - 1;
    if (!xin && !zin) return;
    const int p = z * a.pitch + x, c = z * a.nx + x, pz = a.pitch, cz = a.nx;
    double acc = 0.0;
    if (xin) {
        float su = 0.f, sf = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) {
            su += a.cx[m - 1] * (u[p + m] - u[p - m + 1]);
            sf += a.cx[m - 1] * (fwd[c + m] - fwd[c - m + 1]);
        }
        acc += (double)su * (double)sf;
    }
    if (zin) {
        float su = 0.f, sf = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) {
            su += a.cz[m - 1] * (u[p + m * pz] - u[p - (m - 1) * pz]);
            sf += a.cz[m - 1] * (fwd[c + m * cz] - fwd[c - (m - 1) * cz]);
        }
        acc += (double)su * (double)sf;
    }
    kmu[c] -= acc * scale;
}

void launch_kmu(int R, const StepArgs& a, const float* u, const float* fwd, double* kmu, double scale, hipStream_t s)
{
    const dim3 g((unsigned)((a.nx + kTX - 1) / kTX), (unsigned)((a.nz + 3) / 4)), b(kTX, 4);
    switch (R) {
    case 1: hipLaunchKernelGGL(k_kmu<1>, g, b, 0, s, a, u, fwd, kmu, scale); break;
    case 2: hipLaunchKernelGGL(k_kmu<2>, g, b, 0, s, a, u, fwd, kmu, scale); break;
    case 3: hipLaunchKernelGGL(k_kmu<3>, g, b, 0, s, a, u, fwd, kmu, scale); break;
    default: hipLaunchKernelGGL(k_kmu<4>, g, b, 0, s, a, u, fwd, kmu, scale); break;
    }
}

template <int R>
__global__ __launch_bounds__(256) void k_born_green(StepArgs a, float* sxy, float* szy, const float* dtdmu,
                                                    const float* gin)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)a.nx * a.nz) return;
    const int x = (int)(c % a.nx), z = (int)(c / a.nx);
    const long p = (long)z * a.pitch + x;
    const long nxl = a.nx;
    const float d = dtdmu[c];
    if (x >= R - 1 && x <= a.nx - R - 1) {
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) s += a.cx[m - 1] * (gin[c + m] - gin[c - m + 1]);
        sxy[p] += d * s;
    }
    if (z >= R - 1 && z <= a.nz - R - 1) {
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) s += a.cz[m - 1] * (gin[c + m * nxl] - gin[c - (m - 1) * nxl]);
        szy[p] += d * s;
    }
}

void launch_pair(int R, bool with_u, const PairArgs& p, hipStream_t s)
{
    const dim3 g((unsigned)(kXCD * p.a.tiles_per_xcd)), b(256);
#define FD2D_PAIR(RR)                                                                            \
    case RR:                                                                                     \
        if (with_u) hipLaunchKernelGGL((k_pair<RR, 1>), g, b, 0, s, p);                          \
        else hipLaunchKernelGGL((k_pair<RR, 0>), g, b, 0, s, p);                                 \
        break;
    switch (R) {
        FD2D_PAIR(1)
        FD2D_PAIR(2)
        FD2D_PAIR(3)
        FD2D_PAIR(4)
    }
#undef FD2D_PAIR
}

void launch_step(int R, bool with_u, const StepArgs& a, hipStream_t s)
{
    const dim3 g((unsigned)(kXCD * a.tiles_per_xcd)), b(256);
#define FD2D_TILE(RR)                                                                            \
    case RR:                                                                                     \
        if (with_u) hipLaunchKernelGGL((k_tile<RR, 1>), g, b, 0, s, a);                          \
        else hipLaunchKernelGGL((k_tile<RR, 0>), g, b, 0, s, a);                                 \
        break;
    switch (R) {
        FD2D_TILE(1)
        FD2D_TILE(2)
        FD2D_TILE(3)
        FD2D_TILE(4)
    }
#undef FD2D_TILE
}

void launch_born_green(int R, const StepArgs& a, float* sxy, float* szy, const float* dtdmu, const float* gin,
                       hipStream_t s)
{
    const long n = (long)a.nx * a.nz;
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
    switch (R) {
    case 1: hipLaunchKernelGGL(k_born_green<1>, g, b, 0, s, a, sxy, szy, dtdmu, gin); break;
    case 2: hipLaunchKernelGGL(k_born_green<2>, g, b, 0, s, a, sxy, szy, dtdmu, gin); break;
    case 3: hipLaunchKernelGGL(k_born_green<3>, g, b, 0, s, a, sxy, szy, dtdmu, gin); break;
    default: hipLaunchKernelGGL(k_born_green<4>, g, b, 0, s, a, sxy, szy, dtdmu, gin); break;
    }
}

// ---------------------------------------------------------------------------------------------
// Set-up and element-wise kernels
// ---------------------------------------------------------------------------------------------
__global__ void k_material(const double* mu, const double* rho, float* bv, float* dtmu, int nx, int nz,
                           int pitch, double dt)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)nx * nz) return;
    const int i = (int)(c % nx), j = (int)(c / nx);
    const long p = (long)j * pitch + i;
    bv[p] = (float)(dt / rho[c]);
    dtmu[p] = (float)(dt * mu[c]);
}

void launch_material(const double* mu, const double* rho, float* bv, float* dtmu, int nx, int nz,
                     int pitch, double dt, hipStream_t s)
{
    const long n = (long)nx * nz;
    hipLaunchKernelGGL(k_material, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mu, rho, bv, dtmu,
                       nx, nz, pitch, dt);
}

__global__ void k_scale_to_float(const double* in, float* out, size_t n, double scale)
{
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) out[c] = (float)(scale * in[c]);
}

void launch_scale_to_float(const double* in, float* out, size_t n, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n, scale);
}

__global__ void k_stf_real(float* amp, const double2* stf, int64_t d0, int nq, int nsteps)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nq * nsteps) return;
    const int q = (int)(t % nq);
    const long n = t / nq;
    amp[t] = (float)stf[q + d0 * n].x;
}

void launch_stf_real(float* amp, const double2* stf, int64_t d0, int nq, int nsteps, hipStream_t s)
{
    const long n = (long)nq * nsteps;
    hipLaunchKernelGGL(k_stf_real, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, amp, stf, d0, nq, nsteps);
}

// ---------------------------------------------------------------------------------------------
// Frequency domain
// ---------------------------------------------------------------------------------------------
// The DFT flush and the synthesis of distributed sources run on the matrix cores (spectral.hip);
// these kernels prepare their spectra and finish their results.
__global__ void k_source_spectrum(double2* A, const double2* Z, int conj_in, const double* sd, int n_sd,
                                  const double* I, const double* spec_sum, size_t P, int K)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= P) return;
    for (int k = 0; k < K; ++k) {
        double d = 0.0;
        for (int b = 0; b < n_sd; ++b) d += I[k + (size_t)K * b] * sd[x + P * b];
        const double a = spec_sum[k] * d;
        const double2 z = Z[(size_t)k * P + x];
        A[(size_t)k * P + x] = make_double2(a * z.x, conj_in ? -(a * z.y) : a * z.y);
    }
}

void launch_source_spectrum(double2* A, const double2* Z, int conj_in, const double* sd, int n_sd,
                            const double* I, const double* spec_sum, size_t P, int K, hipStream_t s)
{
    hipLaunchKernelGGL(k_source_spectrum, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, A, Z, conj_in, sd,
                       n_sd, I, spec_sum, P, K);
}

__global__ void k_finalize_dft(double2* out, const double2* acc, size_t n, double scale, int conj)
{
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double2 a = acc[c];
    out[c] = make_double2(scale * a.x, conj ? -(scale * a.y) : scale * a.y);
}

void launch_finalize_dft(double2* out, const double2* acc, size_t n, double scale, int conj, hipStream_t s)
{
    hipLaunchKernelGGL(k_finalize_dft, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, acc, n, scale, conj);
}

__global__ void k_source_kernel(double* Ks, const double2* G, const double2* D, const double* I,
                                const double* spec, size_t P, int K, int nb_cols, double scale)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= P) return;
    for (int b = 0; b < nb_cols; ++b) {
        double s = 0.0;
        for (int k = 0; k < K; ++k) {
            const double2 g = G[(size_t)k * P + x], d = D[(size_t)k * P + x];
            s += I[k + (size_t)K * b] * spec[k] * (g.x * d.x - g.y * d.y);
        }
        Ks[(size_t)b * P + x] = scale * s;
    }
}

void launch_source_kernel(double* Ks, const double2* G, const double2* D, const double* I, const double* spec,
                          size_t P, int K, int nb_cols, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_source_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, Ks, G, D, I, spec, P,
                       K, nb_cols, scale);
}

}  // namespace fd2d
