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
// The structure-kernel product K_mu -= fw_nth * (D u . D u_fwd) of the steps that take it is a fourth
// phase of the same launch (k_kmu after the second step of a two-step launch). The frequency-domain
// work is blocked over time: DFT samples and synthesised source slices are staged as FP32 slabs and
// contracted with the phase tables on the matrix cores in separate launches (spectral.hip).
#include <cfloat>

#include "engine.h"

namespace fd2d {

// ---------------------------------------------------------------------------------------------
// Fused time step
// ---------------------------------------------------------------------------------------------
// One workgroup (kTileThreads / 64 wavefronts) advances one kTX x kTileZ tile by one step:
//   phase 1 issues, in one burst, the loads of sigma over the tile plus the halo that the velocity
//           stencil of the halo needs and the velocity operands, then stores sigma to LDS;
//   phase 2 computes the velocity (and displacement) on the tile plus an (R-1, R) halo in both
//           directions -- the halo is recomputed, bit-identically, by the neighbouring tiles --
//           and writes the outputs of the tile's own nodes;
//   phase 3 updates sigma on the tile.
// Lane l of every wavefront owns column i0 + l; wavefront w owns rows w, w+TW, ... of each region,
// so row offsets are wave-uniform and LDS row offsets fold into instruction immediates. The few
// halo columns (2R-1 of the velocity region, 2(2R-1) of the sigma_xy region) are handled as one
// packed list.
//
// Branch-free by layout: every padded array covers the whole tiles plus their halos, and outside
// the grid the fields, the coefficients dt/rho and dt*mu and the padded taper are zero, so the
// stencils run unguarded over every cell (an outside cell computes 0 and stores 0) and the frames
// of dx_v / dz_v / div_s are 0/1 factors. Only the compact arrays (source slices, snapshots, DFT
// samples, Born inputs) need an in-grid test, folded into a buffer offset past their end for cells
// outside the grid (range-checked buffer loads return 0 there and stores are dropped).
//
// Tiles are dealt to the XCDs in contiguous bands (tile = (g % 8) * tiles_per_xcd + g / 8), so a
// tile's x-neighbour, and a tile row later its z-neighbour, read the shared halo lines from the
// same L2.
//
// MODE 0: Green run (no displacement); 1: with displacement (correlation and adjoint runs);
// 2: with displacement and the structure-kernel product (the adjoint steps that take it). The
// product is a mode of its own so that the other steps carry none of its registers and branches.
__device__ __forceinline__ int node_key(int x, int z) { return ((z + 8) << 16) | (x + 8); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, unsigned off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)off, 0, 0);
}
constexpr unsigned kOutside = 0x80000000u;  // a buffer offset past every compact array

// Loads and stores through global-address-space pointers: the argument structs of a grouped
// launch are read from memory, so the compiler cannot infer the address space of the pointers
// they hold and would otherwise emit flat accesses.
template <class T>
__device__ __forceinline__ T gld(const T* p, int i)
{
    return ((const __attribute__((address_space(1))) T*)p)[i];
}
__device__ __forceinline__ int2 gld(const int2* p, int i)
{
    const long long v = ((const __attribute__((address_space(1))) long long*)p)[i];
    return make_int2((int)v, (int)(v >> 32));
}
template <class T>
__device__ __forceinline__ void gst(T* p, int i, T v)
{
    ((__attribute__((address_space(1))) T*)p)[i] = v;
}

// Addressing shared by the step kernels.
struct Grid {
    int nx, nz, pitch;
    __amdgpu_buffer_rsrc_t rf, rc;
    int o_b, o_m;    // offsets of the dt/rho and mu arrays in the coefficient pool
    float cb, cm;    // added to every coefficient load (the value of a uniform coefficient, else 0)
    __device__ unsigned pofs(int arr, int xx, int zz) const
    {
        return (unsigned)(arr + (zz + kGhost) * pitch + xx + kGhostX) * 4u;
    }
    __device__ float field(int arr, int xx, int zz) const { return bload(rf, pofs(arr, xx, zz)); }
    __device__ float bcoef(int xx, int zz) const { return bload(rc, pofs(o_b, xx, zz)) + cb; }
    __device__ float mcoef(int xx, int zz) const { return bload(rc, pofs(o_m, xx, zz)) + cm; }
    __device__ void put(int arr, int xx, int zz, float v) const { bstore(rf, pofs(arr, xx, zz), v); }
    // byte offset of (xx, zz) in a compact nx x nz array, or past its end outside the grid
    __device__ unsigned cofs(int xx, int zz) const
    {
        return (xx >= 0 && xx < nx && zz >= 0 && zz < nz) ? (unsigned)(zz * nx + xx) * 4u : kOutside;
    }
};

// The grid of a step. A coefficient with the same value at every node (a.binfo: {dt/rho uniform,
// its value, mu uniform, its value}, written at set-up) is taken from binfo, and its array loads
// point past the end of the coefficient pool, where buffer loads return 0 and move no data: a
// homogeneous density or modulus costs no HBM traffic. The values are those the array holds, so the
// results are the same bit for bit (outside the grid the 0/1 frames and the zero taper still give 0).
__device__ __forceinline__ Grid make_grid(const StepArgs& a)
{
    const float* bi = a.binfo;
    const bool ub = bi[0] != 0.f, um = bi[2] != 0.f;
    const int past = (int)(a.cbytes >> 2);
    return Grid{a.nx, a.nz, a.pitch, rsrc(a.fbase, a.fbytes), rsrc(a.cbase, a.cbytes), ub ? past : a.o_bv,
                um ? past : a.o_dtmu, ub ? bi[1] : 0.f, um ? bi[3] : 0.f};
}

template <int R, int MODE>
__device__ __forceinline__ void tile_step(const StepArgs& a, const JobStep& st, int t)
{
    constexpr bool HAS_U = MODE >= 1;
    constexpr int TZ = kTileZ;
    constexpr int TT = kTileThreads, TW = TT / 64;  // threads and wavefronts of the workgroup
    constexpr int WV = kTX + 2 * R - 1;     // velocity region: columns i0-(R-1) .. i0+63+R
    constexpr int HV = TZ + 2 * R - 1;      //                  rows    j0-(R-1) .. j0+TZ-1+R
    constexpr int WSX = WV + 2 * R - 1;     // sigma_xy region: columns i0-(2R-1) .. i0+63+2R-1
    constexpr int HSZ = HV + 2 * R - 1;     // sigma_zy region: rows    j0-(2R-1) .. j0+TZ-1+2R-1
    constexpr int HCV = 2 * R - 1;          // halo columns of the velocity / sigma_zy regions
    constexpr int HCS = 2 * (2 * R - 1);    // halo columns of the sigma_xy region
    __shared__ float sx[HV * WSX + 1];      // + 1: the sink of the packed lists' idle lanes
    __shared__ float sz[HSZ * WV + 1];
    __shared__ float vv[HV * WV + 1];

    if (t >= a.n_tiles) return;
    const int i0 = (t % a.tiles_x) * kTX, j0 = (t / a.tiles_x) * TZ;
    const int nx = a.nx, nz = a.nz;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x = i0 + lane;                 // this lane's own column
    const Grid G = make_grid(a);
    const unsigned cb = a.compact_bytes;
    const bool has_syn = st.syn != nullptr;
    const auto r_syn = rsrc(has_syn ? (const void*)st.syn : (const void*)a.fbase, has_syn ? cb : 0u);
    const bool kp = MODE == 2 && st.kfwd != nullptr;  // this step takes a structure-kernel product

    // ---------------- phase 1
    constexpr int NRV = (HV + TW - 1) / TW, NRZ = (HSZ + TW - 1) / TW;
    constexpr int NHS = (HV * HCS + TT - 1) / TT, NHZ = (HSZ * HCV + TT - 1) / TT, NHV = (HV * HCV + TT - 1) / TT;
    float msx[NRV], msz[NRZ];
    float mv[NRV], mb[NRV], mu[NRV], msy[NRV];
    float hsx[NHS], hsz[NHZ];
    float hv[NHV], hb[NHV], hsy[NHV], ht[NHV], hu[NHV];
#pragma unroll
    for (int k = 0; k < NRV; ++k) {
        const int r = w + TW * k, z = j0 - (R - 1) + r;
        msx[k] = G.field(a.o_sxy_in, x, z);
        mv[k] = G.field(a.o_v_in, x, z);
        mb[k] = G.bcoef(x, z);
        mu[k] = (HAS_U && ((r >= R - 1 && r < R - 1 + TZ) || kp)) ? G.field(a.o_u_in, x, z) : 0.f;
        msy[k] = has_syn ? bload(r_syn, G.cofs(x, z)) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NRZ; ++k) {
        const int r = w + TW * k, z = j0 - (2 * R - 1) + r;
        msz[k] = G.field(a.o_szy_in, x, z);
    }
    // packed halo lists: idle lanes of the last round repeat the last cell and store to the sink
#pragma unroll
    for (int k = 0; k < NHS; ++k) {
        const int idx = min(tid + TT * k, HV * HCS - 1);
        const int r = idx / HCS, h = idx - r * HCS;
        const int c = h < 2 * R - 1 ? h : h + kTX;            // sigma_xy region column
        hsx[k] = G.field(a.o_sxy_in, i0 - (2 * R - 1) + c, j0 - (R - 1) + r);
    }
#pragma unroll
    for (int k = 0; k < NHZ; ++k) {
        const int idx = min(tid + TT * k, HSZ * HCV - 1);
        const int r = idx / HCV, h = idx - r * HCV;
        const int c = h < R - 1 ? h : h + kTX;                // velocity-region column
        hsz[k] = G.field(a.o_szy_in, i0 - (R - 1) + c, j0 - (2 * R - 1) + r);
    }
#pragma unroll
    for (int k = 0; k < NHV; ++k) {
        const int idx = min(tid + TT * k, HV * HCV - 1);
        const int r = idx / HCV, h = idx - r * HCV;
        const int c = h < R - 1 ? h : h + kTX;
        const int xx = i0 - (R - 1) + c, zz = j0 - (R - 1) + r;
        hv[k] = G.field(a.o_v_in, xx, zz);
        hb[k] = G.bcoef(xx, zz);
        hsy[k] = has_syn ? bload(r_syn, G.cofs(xx, zz)) : 0.f;
        ht[k] = gld(a.axp, xx + kGhostX) * gld(a.azp, zz + kGhost);
        hu[k] = (HAS_U && kp) ? G.field(a.o_u_in, xx, zz) : 0.f;
    }
    // the forward snapshot of a product step over the velocity region, and the K_mu values of the
    // tile's own nodes (used in phase 4; issued here so their latency hides behind phases 1-3)
    constexpr int NT = TZ / TW;
    float fm[NRV], fh[NHV];
    double kold[NT];
    if (kp) {
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const unsigned co = G.cofs(x, j0 + w + TW * k);
            kold[k] = co != kOutside ? gld(a.kacc, (int)(co / 4u)) : 0.0;
        }
        const auto r_fw = rsrc(st.kfwd, cb);
#pragma unroll
        for (int k = 0; k < NRV; ++k) fm[k] = bload(r_fw, G.cofs(x, j0 - (R - 1) + w + TW * k));
#pragma unroll
        for (int k = 0; k < NHV; ++k) {
            const int idx = min(tid + TT * k, HV * HCV - 1);
            const int r = idx / HCV, h = idx - r * HCV;
            fh[k] = bload(r_fw, G.cofs(i0 - (R - 1) + (h < R - 1 ? h : h + kTX), j0 - (R - 1) + r));
        }
    }
    // LDS stores
#pragma unroll
    for (int k = 0; k < NRV; ++k) {
        const int r = w + TW * k;
        if (r < HV) sx[r * WSX + lane + 2 * R - 1] = msx[k];
    }
#pragma unroll
    for (int k = 0; k < NRZ; ++k) {
        const int r = w + TW * k;
        if (r < HSZ) sz[r * WV + lane + R - 1] = msz[k];
    }
#pragma unroll
    for (int k = 0; k < NHS; ++k) {
        const int idx = tid + TT * k;
        const int r = idx / HCS, h = idx - r * HCS;
        sx[idx < HV * HCS ? r * WSX + (h < 2 * R - 1 ? h : h + kTX) : HV * WSX] = hsx[k];
    }
#pragma unroll
    for (int k = 0; k < NHZ; ++k) {
        const int idx = tid + TT * k;
        const int r = idx / HCV, h = idx - r * HCV;
        sz[idx < HSZ * HCV ? r * WV + (h < R - 1 ? h : h + kTX) : HSZ * WV] = hsz[k];
    }
    __syncthreads();

    // ---------------- phase 2
    const int e0 = a.ps_off ? gld(a.ps_off, t) : 0, e1 = a.ps_off ? gld(a.ps_off, t + 1) : 0;
    const int f0 = a.rec_off ? gld(a.rec_off, t) : 0, f1 = a.rec_off ? gld(a.rec_off, t + 1) : 0;
    // div_s at region cell (r, c) = node (xx, zz) from the staged sigma, with the frame factors
    // fx (x-part, R <= x <= nx-R-1) and fz (z-part, R <= z <= nz-R-1), plus the point sources
    auto divs = [&](int xx, int zz, int r, int c, float fx, float fz) {
        const float* row = sx + r * WSX + c + R;
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) s += a.cx[m - 1] * (row[m - 1] - row[-m]);
        const float* col = sz + (r + R) * WV + c;
        float q = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) q += a.cz[m - 1] * (col[(m - 1) * WV] - col[-m * WV]);
        float ds = s * fx + q * fz;
        if (e1 > e0) {
            const int key = node_key(xx, zz);
            for (int e = e0; e < e1; ++e) {
                const int2 en = gld(a.ps_ent, e);
                if (en.x == key) ds += st.ps_amp ? gld(st.ps_amp, en.y) : st.ps_const;
            }
        }
        return ds;
    };
    auto frame = [](int v, int lo, int hi) { return (v >= lo && v <= hi) ? 1.f : 0.f; };
    const float fxl = frame(x, R, nx - R - 1);
    const float axl = gld(a.axp, x + kGhostX);
#pragma unroll
    for (int k = 0; k < NRV; ++k) {
        const int r = w + TW * k;
        if (r >= HV) break;
        const int z = j0 - (R - 1) + r, c = lane + R - 1;
        const float ds = divs(x, z, r, c, fxl, frame(z, R, nz - R - 1)) + msy[k];
        const float vn = (mv[k] + mb[k] * ds) * (axl * gld(a.azp, z + kGhost));
        if (r >= R - 1 && r < R - 1 + TZ) {  // own node: outputs (zeros outside the grid)
            G.put(a.o_v_out, x, z, vn);
            float out = vn;
            const unsigned co = G.cofs(x, z);
            if (HAS_U) {
                const float un = mu[k] + vn * a.dt;
                G.put(a.o_u_out, x, z, un);
                if (st.snap) bstore(rsrc(st.snap, cb), co, mu[k]);
                out = un;
            } else if (st.snap) {
                bstore(rsrc(st.snap, cb), co, mv[k]);
            }
            if (st.dsamp) bstore(rsrc(st.dsamp, cb), co, out);
            if (f1 > f0 && st.seis_col) {
                const int key = node_key(x, z);
                for (int e = f0; e < f1; ++e) {
                    const int2 en = gld(a.rec_ent, e);
                    if (en.x == key) gst(st.seis_col, en.y, (double)out);
                }
            }
        }
        vv[r * WV + c] = vn;
    }
#pragma unroll
    for (int k = 0; k < NHV; ++k) {
        const int idx = min(tid + TT * k, HV * HCV - 1);
        const int r = idx / HCV, h = idx - r * HCV;
        const int c = h < R - 1 ? h : h + kTX;
        const int xx = i0 - (R - 1) + c, zz = j0 - (R - 1) + r;
        const float ds = divs(xx, zz, r, c, frame(xx, R, nx - R - 1), frame(zz, R, nz - R - 1)) + hsy[k];
        const float vn = (hv[k] + hb[k] * ds) * ht[k];
        vv[tid + TT * k < HV * HCV ? r * WV + c : HV * WV] = vn;
    }
    __syncthreads();

    // ---------------- phase 3
    float rm[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) rm[k] = G.mcoef(x, j0 + w + TW * k);
    const float fxs = frame(x, R - 1, nx - R - 1);  // dx_v frame
    const int c = lane + R - 1;
    const bool born = st.born_dmu != nullptr;
    // born_in1 is null for the Green run's Born term, which has one snapshot: its loads return 0
    const bool born1 = born && st.born_in1 != nullptr;
    const auto r_b0 = rsrc(born ? (const void*)st.born_in0 : (const void*)a.fbase, born ? cb : 0u);
    const auto r_b1 = rsrc(born1 ? (const void*)st.born_in1 : (const void*)a.fbase, born1 ? cb : 0u);
    const auto r_bd = rsrc(born ? (const void*)st.born_dmu : (const void*)a.fbase, born ? cb : 0u);
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const int lz = w + TW * k, z = j0 + lz;
        const int r = lz + R - 1;
        const float fzs = frame(z, R - 1, nz - R - 1);  // dz_v frame
        const float dm = rm[k];
        const float* vrow = vv + r * WV + c;
        float ex = 0.f, ez = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) ex += a.cx[m - 1] * (vrow[m] - vrow[1 - m]);
#pragma unroll
        for (int m = 1; m <= R; ++m) ez += a.cz[m - 1] * (vrow[m * WV] - vrow[(1 - m) * WV]);
        float snx = sx[r * WSX + c + R] + dm * (ex * fxs);
        float snz = sz[(r + R) * WV + c] + dm * (ez * fzs);
        if (born) {
            // Born terms: - dmu * D(C_in(n) - C_in(n+1)) (FP32 snapshots, as the reference); in the
            // Green run + dt dmu D G_in(n+1), the next step's term (born_dmu = -dt dmu, no C_in(n+1))
            const unsigned co = G.cofs(x, z);
            auto bd = [&](unsigned o) { return bload(r_b0, o) - bload(r_b1, o); };
            float bx = 0.f, bz = 0.f;
            if (co != kOutside) {
                const unsigned nx4 = (unsigned)nx * 4u;
#pragma unroll
                for (int m = 1; m <= R; ++m) bx += a.cx[m - 1] * (bd(co + 4u * m) - bd(co - 4u * (m - 1)));
#pragma unroll
                for (int m = 1; m <= R; ++m) bz += a.cz[m - 1] * (bd(co + nx4 * m) - bd(co - nx4 * (m - 1)));
            }
            const float dmu = bload(r_bd, co);
            snx -= dmu * (bx * fxs);
            snz -= dmu * (bz * fzs);
        }
        G.put(a.o_sxy_out, x, z, snx);
        G.put(a.o_szy_out, x, z, snz);
    }

    // ---------------- phase 4 (steps with a structure-kernel product)
    // K_mu -= kscale * (D u . D u_fwd) with the displacement of this step, the forward snapshot kfwd
    // and the dx_v / dz_v stencils on their frames (run_noise_adjoint.m:324-331). The displacement
    // (in vv, in place) and the snapshot (in sx, row pitch WV) are staged in LDS over the velocity
    // region, which covers the (R-1, R) halo of the stencils, so no other tile's output is read;
    // the sums are those of k_kmu, term for term.
    if (kp) {
        const bool xin = x >= R - 1 && x <= nx - R - 1;
        __syncthreads();  // phase 3 has read vv, sx and sz
#pragma unroll
        for (int k = 0; k < NRV; ++k) {
            const int r = w + TW * k;
            if (r >= HV) break;
            float* cell = vv + r * WV + lane + R - 1;
            *cell = mu[k] + *cell * a.dt;
            sx[r * WV + lane + R - 1] = fm[k];
        }
#pragma unroll
        for (int k = 0; k < NHV; ++k) {
            const int idx = min(tid + TT * k, HV * HCV - 1);
            const int r = idx / HCV, h = idx - r * HCV;
            const int o = tid + TT * k < HV * HCV ? r * WV + (h < R - 1 ? h : h + kTX) : HV * WV;
            vv[o] = hu[k] + vv[o] * a.dt;
            sx[o] = fh[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int lz = w + TW * k, z = j0 + lz;
            const unsigned co = G.cofs(x, z);
            const bool zin = z >= R - 1 && z <= nz - R - 1;
            if (co == kOutside || (!xin && !zin)) continue;
            const int o = (lz + R - 1) * WV + c;
            const float* urow = vv + o;
            const float* frow = sx + o;
            double acc = 0.0;
            if (xin) {
                float su = 0.f, sf = 0.f;
#pragma unroll
                for (int m = 1; m <= R; ++m) {
                    su += a.cx[m - 1] * (urow[m] - urow[1 - m]);
                    sf += a.cx[m - 1] * (frow[m] - frow[1 - m]);
                }
                acc += (double)su * (double)sf;
            }
            if (zin) {
                float su = 0.f, sf = 0.f;
#pragma unroll
                for (int m = 1; m <= R; ++m) {
                    su += a.cz[m - 1] * (urow[m * WV] - urow[(1 - m) * WV]);
                    sf += a.cz[m - 1] * (frow[m * WV] - frow[(1 - m) * WV]);
                }
                acc += (double)su * (double)sf;
            }
            gst(a.kacc, (int)(co / 4u), kold[k] - acc * a.kscale);
        }
    }
}

template <int R, int MODE>
__global__ __launch_bounds__(kTileThreads) void k_tile(StepArgs a)
{
    const int g = blockIdx.x;
    const JobStep st{a.ps_amp, a.syn, a.snap, a.dsamp, a.seis_col, a.born_dmu, a.born_in0, a.born_in1, a.kfwd,
                     a.ps_const, 0, 0};
    tile_step<R, MODE>(a, st, (g % kXCD) * a.tiles_per_xcd + g / kXCD);
}

template <int R, int MODE>
__global__ __launch_bounds__(kTileThreads) void k_tile_group(GroupArgs G)
{
    const int g = blockIdx.x;
    const int T = (g % kXCD) * G.tiles_per_xcd + g / kXCD;
    if (T >= G.total_tiles) return;
    int j = 0;
    while (j + 1 < G.n_jobs && G.tile_start[j + 1] <= T) ++j;  // wave-uniform scan
    j = __builtin_amdgcn_readfirstlane(j);
    // by-value copies: the step body then reads its arguments from registers, not through
    // pointers that the compiler must assume its own stores may alias
    const JobStep st = G.step[j];
    const StepArgs a = G.jobs[2 * st.slot + st.parity];
    tile_step<R, MODE>(a, st, T - G.tile_start[j]);
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
// the rows w, w+PW, ...; halo columns are packed lists over the kPairThreads threads.
template <int R, int MODE>
__device__ __forceinline__ void pair_step(const StepArgs& a, const PairStep& ps, int t)
{
    constexpr bool HAS_U = MODE >= 1;
    constexpr int TZ = kTileZ;
    constexpr int PT = kPairThreads, PW = PT / 64;  // threads and wavefronts of the workgroup
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
    __shared__ float SX[H1 * WSX + 1];       // + 1: the sink of the packed lists' idle lanes
    __shared__ float SZ[HSZ * W1 + 1];
    __shared__ float VV[H1 * W1 + 1];

    if (t >= a.n_tiles) return;
    const int i0 = (t % a.tiles_x) * kTX, j0 = (t / a.tiles_x) * TZ;
    const int nx = a.nx, nz = a.nz;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x = i0 + lane;
    const Grid G = make_grid(a);
    const unsigned cb = a.compact_bytes;
    // the synthesised source of the pair belongs to step n (ps.syn) or to step n+1 (syn2)
    const float* syn = ps.syn ? ps.syn : ps.syn2;
    const bool syn1 = ps.syn != nullptr, has_syn = syn != nullptr;
    const auto r_syn = rsrc(has_syn ? (const void*)syn : (const void*)a.fbase, has_syn ? cb : 0u);
    auto frame = [](int v, int lo, int hi) { return (v >= lo && v <= hi) ? 1.f : 0.f; };

    // ---------------- phase A: stage sigma, load the step-n velocity operands of V1
    constexpr int NM1 = (H1 + PW - 1) / PW, NH1 = (H1 * HC1 + PT - 1) / PT;
    constexpr int NMZ = (HSZ + PW - 1) / PW;
    constexpr int NHX = (H1 * HCS + PT - 1) / PT, NHZ = (HSZ * HC1 + PT - 1) / PT;
    float mv[NM1], mb[NM1], msy[NM1], mu[NM1];
    float hv[NH1], hb[NH1], ht[NH1], hsy[NH1];
    {
        float sxm[NM1], sxh[NHX], szm[NMZ], szh[NHZ];
#pragma unroll
        for (int k = 0; k < NM1; ++k) sxm[k] = G.field(a.o_sxy_in, x, j0 - L1 + w + PW * k);
#pragma unroll
        for (int k = 0; k < NHX; ++k) {
            const int idx = min(tid + PT * k, H1 * HCS - 1);
            const int r = idx / HCS, h = idx - r * HCS;
            const int c = h < LS ? h : h + kTX;
            sxh[k] = G.field(a.o_sxy_in, i0 - LS + c, j0 - L1 + r);
        }
#pragma unroll
        for (int k = 0; k < NMZ; ++k) szm[k] = G.field(a.o_szy_in, x, j0 - LS + w + PW * k);
#pragma unroll
        for (int k = 0; k < NHZ; ++k) {
            const int idx = min(tid + PT * k, HSZ * HC1 - 1);
            const int r = idx / HC1, h = idx - r * HC1;
            const int c = h < L1 ? h : h + kTX;
            szh[k] = G.field(a.o_szy_in, i0 - L1 + c, j0 - LS + r);
        }
#pragma unroll
        for (int k = 0; k < NM1; ++k) {
            const int r = w + PW * k, z = j0 - L1 + r;
            mv[k] = G.field(a.o_v_in, x, z);
            mb[k] = G.bcoef(x, z);
            mu[k] = (HAS_U && r >= L1 && r < L1 + TZ) ? G.field(a.o_u_in, x, z) : 0.f;
            msy[k] = has_syn ? bload(r_syn, G.cofs(x, z)) : 0.f;
        }
#pragma unroll
        for (int k = 0; k < NH1; ++k) {
            const int idx = min(tid + PT * k, H1 * HC1 - 1);
            const int r = idx / HC1, h = idx - r * HC1;
            const int c = h < L1 ? h : h + kTX;
            const int xx = i0 - L1 + c, zz = j0 - L1 + r;
            hv[k] = G.field(a.o_v_in, xx, zz);
            hb[k] = G.bcoef(xx, zz);
            hsy[k] = has_syn ? bload(r_syn, G.cofs(xx, zz)) : 0.f;
            ht[k] = gld(a.axp, xx + kGhostX) * gld(a.azp, zz + kGhost);
        }
#pragma unroll
        for (int k = 0; k < NM1; ++k) {
            const int r = w + PW * k;
            if (r < H1) SX[r * WSX + lane + LS] = sxm[k];
        }
#pragma unroll
        for (int k = 0; k < NHX; ++k) {
            const int idx = tid + PT * k;
            const int r = idx / HCS, h = idx - r * HCS;
            SX[idx < H1 * HCS ? r * WSX + (h < LS ? h : h + kTX) : H1 * WSX] = sxh[k];
        }
#pragma unroll
        for (int k = 0; k < NMZ; ++k) {
            const int r = w + PW * k;
            if (r < HSZ) SZ[r * W1 + lane + L1] = szm[k];
        }
#pragma unroll
        for (int k = 0; k < NHZ; ++k) {
            const int idx = tid + PT * k;
            const int r = idx / HC1, h = idx - r * HC1;
            SZ[idx < HSZ * HC1 ? r * W1 + (h < L1 ? h : h + kTX) : HSZ * W1] = szh[k];
        }
    }
    __syncthreads();

    // ---------------- phase B: step-n velocity on V1
    const int e0 = a.ps_off ? gld(a.ps_off, t) : 0, e1 = a.ps_off ? gld(a.ps_off, t + 1) : 0;
    const int f0 = a.rec_off ? gld(a.rec_off, t) : 0, f1 = a.rec_off ? gld(a.rec_off, t + 1) : 0;
    auto divs = [&](int xx, int zz, int r, int c, float fx, float fz, const float* amp, float cst) {
        const float* row = SX + r * WSX + c + R;
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) s += a.cx[m - 1] * (row[m - 1] - row[-m]);
        const float* col = SZ + (r + R) * W1 + c;
        float q = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) q += a.cz[m - 1] * (col[(m - 1) * W1] - col[-m * W1]);
        float ds = s * fx + q * fz;
        if (e1 > e0) {
            const int key = node_key(xx, zz);
            for (int e = e0; e < e1; ++e) {
                const int2 en = gld(a.ps_ent, e);
                if (en.x == key) ds += amp ? gld(amp, en.y) : cst;
            }
        }
        return ds;
    };
    auto gather = [&](int xx, int zz, float val, double* col) {
        if (f1 > f0 && col) {
            const int key = node_key(xx, zz);
            for (int e = f0; e < f1; ++e) {
                const int2 en = gld(a.rec_ent, e);
                if (en.x == key) gst(col, en.y, (double)val);
            }
        }
    };
    // dt*mu of the sigma updates, prefetched: main columns over the sigma_zy rows (which contain
    // the sigma_xy rows and the tile), and the two packed halo lists
    constexpr int NMD = (HS1Z + PW - 1) / PW;
    constexpr int NDX = (HS1X * HCX1 + PT - 1) / PT, NDZ = (HS1Z * HCZ1 + PT - 1) / PT;
    float dmm[NMD], dmx[NDX], dmz[NDZ];
#pragma unroll
    for (int k = 0; k < NMD; ++k) dmm[k] = G.mcoef(x, j0 - (2 * R - 1) + w + PW * k);
#pragma unroll
    for (int k = 0; k < NDX; ++k) {
        const int idx = min(tid + PT * k, HS1X * HCX1 - 1);
        const int r = idx / HCX1, h = idx - r * HCX1;
        const int xx = h < 2 * R - 1 ? i0 - (2 * R - 1) + h : i0 + kTX + (h - (2 * R - 1));
        dmx[k] = G.mcoef(xx, j0 - (R - 1) + r);
    }
#pragma unroll
    for (int k = 0; k < NDZ; ++k) {
        const int idx = min(tid + PT * k, HS1Z * HCZ1 - 1);
        const int r = idx / HCZ1, h = idx - r * HCZ1;
        const int xx = h < R - 1 ? i0 - (R - 1) + h : i0 + kTX + (h - (R - 1));
        dmz[k] = G.mcoef(xx, j0 - (2 * R - 1) + r);
    }

    const float fxl = frame(x, R, nx - R - 1);
    const float axl = gld(a.axp, x + kGhostX);
    float un1[NM1];  // displacement after step n at the tile's own rows
#pragma unroll
    for (int k = 0; k < NM1; ++k) {
        const int r = w + PW * k;
        un1[k] = 0.f;
        if (r >= H1) break;
        const int z = j0 - L1 + r, c = lane + L1;
        const float ds = divs(x, z, r, c, fxl, frame(z, R, nz - R - 1), ps.ps_amp, ps.ps_const) + (syn1 ? msy[k] : 0.f);
        const float vn = (mv[k] + mb[k] * ds) * (axl * gld(a.azp, z + kGhost));
        if (r >= L1 && r < L1 + TZ) {  // own node: outputs of step n
            const float un = HAS_U ? mu[k] + vn * a.dt : 0.f;
            un1[k] = un;
            const unsigned co = G.cofs(x, z);
            float out = vn;
            if (HAS_U) {
                if (ps.snap) bstore(rsrc(ps.snap, cb), co, mu[k]);
                out = un;
            } else if (ps.snap) {
                bstore(rsrc(ps.snap, cb), co, mv[k]);
            }
            if (ps.dsamp) bstore(rsrc(ps.dsamp, cb), co, out);
            gather(x, z, out, ps.seis_col);
        }
        VV[r * W1 + c] = vn;
    }
#pragma unroll
    for (int k = 0; k < NH1; ++k) {
        const int idx = min(tid + PT * k, H1 * HC1 - 1);
        const int r = idx / HC1, h = idx - r * HC1;
        const int c = h < L1 ? h : h + kTX;
        const int xx = i0 - L1 + c, zz = j0 - L1 + r;
        const float ds = divs(xx, zz, r, c, frame(xx, R, nx - R - 1), frame(zz, R, nz - R - 1), ps.ps_amp, ps.ps_const) +
                         (syn1 ? hsy[k] : 0.f);
        VV[tid + PT * k < H1 * HC1 ? r * W1 + c : H1 * W1] = (hv[k] + hb[k] * ds) * ht[k];
    }
    __syncthreads();

    // ---------------- phase C: step-n sigma, in place, where step n+1 needs it
    const float fxs = frame(x, R - 1, nx - R - 1);  // dx_v frame
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
        const int q = w + PW * k;
        if (q >= HS1Z) break;
        const int z = j0 - (2 * R - 1) + q;
        const int rv = z - (j0 - L1), c = lane + L1;
        SZ[(rv + R) * W1 + c] += dmm[k] * (dzv(rv, c) * frame(z, R - 1, nz - R - 1));
        if (rv >= V2R0 && rv <= V2R1) SX[rv * WSX + c + R] += dmm[k] * (dxv(rv, c) * fxs);
    }
#pragma unroll
    for (int k = 0; k < NDX; ++k) {
        const int idx = min(tid + PT * k, HS1X * HCX1 - 1);
        const int r = idx / HCX1, h = idx - r * HCX1;
        const int xx = h < 2 * R - 1 ? i0 - (2 * R - 1) + h : i0 + kTX + (h - (2 * R - 1));
        const int zz = j0 - (R - 1) + r;
        const int rv = zz - (j0 - L1), cv = xx - (i0 - L1);
        const float upd = dmx[k] * (dxv(rv, cv) * frame(xx, R - 1, nx - R - 1));
        if (tid + PT * k < HS1X * HCX1) SX[rv * WSX + cv + R] += upd;
    }
#pragma unroll
    for (int k = 0; k < NDZ; ++k) {
        const int idx = min(tid + PT * k, HS1Z * HCZ1 - 1);
        const int r = idx / HCZ1, h = idx - r * HCZ1;
        const int xx = h < R - 1 ? i0 - (R - 1) + h : i0 + kTX + (h - (R - 1));
        const int zz = j0 - (2 * R - 1) + r;
        const int rv = zz - (j0 - L1), cv = xx - (i0 - L1);
        const float upd = dmz[k] * (dzv(rv, cv) * frame(zz, R - 1, nz - R - 1));
        if (tid + PT * k < HS1Z * HCZ1) SZ[(rv + R) * W1 + cv] += upd;
    }
    __syncthreads();

    // ---------------- phase D: step-(n+1) velocity on V2, in place
#pragma unroll
    for (int k = 0; k < NM1; ++k) {
        const int r = w + PW * k;
        if (r >= H1) break;
        if (r < V2R0 || r > V2R1) continue;
        const int z = j0 - L1 + r, c = lane + L1;
        const float v1 = VV[r * W1 + c];
        const float ds = divs(x, z, r, c, fxl, frame(z, R, nz - R - 1), ps.ps_amp2, ps.ps_const2) + (syn1 ? 0.f : msy[k]);
        const float vn = (v1 + mb[k] * ds) * (axl * gld(a.azp, z + kGhost));
        if (r >= L1 && r < L1 + TZ) {  // own node: outputs of step n+1 and the new state
            G.put(a.o_v_out, x, z, vn);
            const unsigned co = G.cofs(x, z);
            float out = vn;
            if (HAS_U) {
                const float un = un1[k] + vn * a.dt;
                G.put(a.o_u_out, x, z, un);
                if (ps.snap2) bstore(rsrc(ps.snap2, cb), co, un1[k]);
                out = un;
            } else if (ps.snap2) {
                bstore(rsrc(ps.snap2, cb), co, v1);
            }
            if (ps.dsamp2) bstore(rsrc(ps.dsamp2, cb), co, out);
            gather(x, z, out, ps.seis_col2);
        }
        VV[r * W1 + c] = vn;
    }
#pragma unroll
    for (int k = 0; k < NH1; ++k) {
        const int idx = min(tid + PT * k, H1 * HC1 - 1);
        const int r = idx / HC1, h = idx - r * HC1;
        const int c = h < L1 ? h : h + kTX;
        const bool in2 = tid + PT * k < H1 * HC1 && r >= V2R0 && r <= V2R1 && c >= V2C0 && c <= V2C1;
        const int xx = i0 - L1 + c, zz = j0 - L1 + r;
        const float ds = divs(xx, zz, r, c, frame(xx, R, nx - R - 1), frame(zz, R, nz - R - 1), ps.ps_amp2,
                              ps.ps_const2) + (syn1 ? 0.f : hsy[k]);
        const float vn = (VV[r * W1 + c] + hb[k] * ds) * ht[k];
        VV[in2 ? r * W1 + c : H1 * W1] = vn;
    }
    __syncthreads();

    // ---------------- phase E: step-(n+1) sigma on the tile
#pragma unroll
    for (int k = 0; k < NMD; ++k) {
        const int q = w + PW * k;
        if (q >= HS1Z) break;
        const int z = j0 - (2 * R - 1) + q;
        if (z < j0 || z >= j0 + TZ) continue;
        const int rv = z - (j0 - L1), c = lane + L1;
        G.put(a.o_sxy_out, x, z, SX[rv * WSX + c + R] + dmm[k] * (dxv(rv, c) * fxs));
        G.put(a.o_szy_out, x, z, SZ[(rv + R) * W1 + c] + dmm[k] * (dzv(rv, c) * frame(z, R - 1, nz - R - 1)));
    }
}


template <int R, int MODE>
__global__ __launch_bounds__(kPairThreads) void k_pair(PairArgs pa)
{
    const int g = blockIdx.x;
    const PairStep ps{pa.a.ps_amp, pa.a.syn, pa.a.snap, pa.a.dsamp, pa.a.seis_col, pa.ps_amp2, pa.syn2, pa.snap2,
                      pa.dsamp2, pa.seis_col2, pa.a.ps_const, pa.ps_const2, 0, 0};
    pair_step<R, MODE>(pa.a, ps, (g % kXCD) * pa.a.tiles_per_xcd + g / kXCD);
}

template <int R, int MODE>
__global__ __launch_bounds__(kPairThreads) void k_pair_group(PairGroupArgs G)
{
    const int g = blockIdx.x;
    const int T = (g % kXCD) * G.tiles_per_xcd + g / kXCD;
    if (T >= G.total_tiles) return;
    int j = 0;
    while (j + 1 < G.n_jobs && G.tile_start[j + 1] <= T) ++j;  // wave-uniform scan
    j = __builtin_amdgcn_readfirstlane(j);
    const PairStep ps = G.step[j];
    const StepArgs a = G.jobs[2 * ps.slot + ps.parity];
    pair_step<R, MODE>(a, ps, T - G.tile_start[j]);
}

// Structure-kernel product of run_noise_adjoint.m:324-331 after a step that stores it:
// ---------------------------------------------------------------------------------------------
// Displacement-form step (correlation and adjoint runs without Born terms, FD orders 6 and 8)
// ---------------------------------------------------------------------------------------------
// In the correlation and adjoint runs sigma and u both start at zero and are advanced by the same
// velocity, sigma_n = sigma_{n-1} + dt mu D v_n and u_n = u_{n-1} + dt v_n
// (code/run_forward2_correlation.m:220-242, code/run_noise_adjoint.m:277-295), so sigma_n = mu D u_n
// at every step, D being linear. This step keeps v (FP32) and u (FP64) only and evaluates sigma from
// u where the divergence needs it: 32 B per node and step instead of the 40 B of v, sigma_xy,
// sigma_zy and u, and sigma carries the rounding of one FP32 evaluation from the FP64 differences of
// u instead of that of a running FP32 sum. One 512-thread workgroup per 64 x 32 tile:
//   phase 1 stages u (FP64) over the tile rows with a (2R-1)-column halo and over the tile columns
//           with a (2R-1)-row halo (the corners are not needed), and loads v, dt/rho, mu, the source
//           slice and, for a product, the forward snapshot (tile + (R-1, R) halo) and K_mu;
//   phase 2 sigma_xy = mu dx_v(u) on the tile rows over columns i0-R .. i0+63+R-1 and
//           sigma_zy = mu dz_v(u) on the tile columns over rows j0-R .. j0+31+R-1 (LDS), with the
//           zero frames of dx_v / dz_v;
//   phase 3 the divergence (frames of div_s), point sources and source slice, v_n with the taper,
//           u_n = u_{n-1} + dt v_n in FP64, the snapshot (u before the update), DFT sample and
//           receivers (u after it);
//   phase 4 (MODE 1) the structure-kernel product of the PREVIOUS step, K_mu -= f (D u_{n-1} . D fwd)
//           (run_noise_adjoint.m:324-331): a product step's displacement is the next step's staged
//           input, so the product needs no halo work of its own (the last step's product, if any,
//           runs in k_ukmu).
__device__ __forceinline__ double bload8(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bstore8(__amdgpu_buffer_rsrc_t r, unsigned off, double v)
{
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)off, 0, 0);
}
// byte offset of (xx, zz) in a padded FP64 array starting at FP32 element `arr` of the field pool
__device__ __forceinline__ unsigned pofs8(int arr, int pitch, int xx, int zz)
{
    return (unsigned)arr * 4u + (unsigned)((zz + kGhost) * pitch + xx + kGhostX) * 8u;
}

template <int R, int MODE>
__device__ __forceinline__ void utile_step(const StepArgs& a, const JobStep& st, int t)
{
    constexpr int TZ = kTileZ;
    constexpr int TT = kTileThreads, TW = TT / 64;
    constexpr int H = 2 * R - 1;                     // u halo
    constexpr int WU = kTX + 2 * H, HU = TZ + 2 * H;  // u region: columns i0-H .., rows j0-H ..
    constexpr int WX = kTX + 2 * R;                  // sigma_xy: columns i0-R .. i0+63+R-1, tile rows
    constexpr int HZ = TZ + 2 * R;                   // sigma_zy: rows j0-R .. j0+TZ-1+R-1, tile columns
    constexpr int WF = kTX + 2 * R - 1, HF = TZ + 2 * R - 1;  // forward snapshot: (R-1, R) halo
    constexpr int HCF = 2 * R - 1;                   // its halo columns
    __shared__ double uu[HU * WU + 1];               // + 1: the sink of the packed lists' idle lanes
    // sigma_xy and sigma_zy, then (phase 4) the forward snapshot in the same space, so a product
    // step holds no more LDS than a plain one (three workgroups per CU)
    __shared__ float sg[TZ * WX + 1 + HZ * kTX];
    float* const sxy = sg;
    float* const szy = sg + TZ * WX + 1;
    float* const ff = sg;
    static_assert(HF * WF + 1 <= TZ * WX + 1 + HZ * kTX, "the forward snapshot must fit the sigma space");

    if (t >= a.n_tiles) return;
    const int i0 = (t % a.tiles_x) * kTX, j0 = (t / a.tiles_x) * TZ;
    const int nx = a.nx, nz = a.nz, pitch = a.pitch;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x = i0 + lane;
    const Grid G = make_grid(a);
    const unsigned cb = a.compact_bytes;
    const bool has_syn = st.syn != nullptr;
    const auto r_syn = rsrc(has_syn ? (const void*)st.syn : (const void*)a.fbase, has_syn ? cb : 0u);
    const bool kp = MODE == 1 && st.kfwd != nullptr;

    // ---------------- phase 1
    constexpr int NRU = (HU + TW - 1) / TW;             // u column block rows per wavefront
    constexpr int NHU = (TZ * 2 * H + TT - 1) / TT;     // u halo columns on the tile rows (packed)
    constexpr int NRZ = (HZ + TW - 1) / TW;             // mu over the sigma_zy rows (tile columns)
    constexpr int NHM = (TZ * 2 * R + TT - 1) / TT;     // mu at the sigma_xy halo columns (packed)
    constexpr int NT = TZ / TW;
    double cu[NRU], hu[NHU];
    float mz[NRZ], hm[NHM];
    float mv[NT], mb[NT], msy[NT];
#pragma unroll
    for (int k = 0; k < NRU; ++k) cu[k] = bload8(G.rf, pofs8(a.o_u_in, pitch, x, j0 - H + w + TW * k));
#pragma unroll
    for (int k = 0; k < NHU; ++k) {
        const int idx = min(tid + TT * k, TZ * 2 * H - 1);
        const int r = idx / (2 * H), h = idx - r * 2 * H;
        const int c = h < H ? h : h + kTX;  // u region column
        hu[k] = bload8(G.rf, pofs8(a.o_u_in, pitch, i0 - H + c, j0 + r));
    }
#pragma unroll
    for (int k = 0; k < NRZ; ++k) mz[k] = G.mcoef(x, j0 - R + w + TW * k);
#pragma unroll
    for (int k = 0; k < NHM; ++k) {
        const int idx = min(tid + TT * k, TZ * 2 * R - 1);
        const int r = idx / (2 * R), h = idx - r * 2 * R;
        hm[k] = G.mcoef(i0 - R + (h < R ? h : h + kTX), j0 + r);
    }
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const int z = j0 + w + TW * k;
        mv[k] = G.field(a.o_v_in, x, z);
        mb[k] = G.bcoef(x, z);
        msy[k] = has_syn ? bload(r_syn, G.cofs(x, z)) : 0.f;
    }
    constexpr int NRF = (HF + TW - 1) / TW, NHF = (HF * HCF + TT - 1) / TT;
    float fm[MODE == 1 ? NRF : 1], fh[MODE == 1 ? NHF : 1];
    double kold[MODE == 1 ? NT : 1];
    if (kp) {
        const auto r_fw = rsrc(st.kfwd, cb);
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const unsigned co = G.cofs(x, j0 + w + TW * k);
            kold[k] = co != kOutside ? gld(a.kacc, (int)(co / 4u)) : 0.0;
        }
#pragma unroll
        for (int k = 0; k < NRF; ++k) fm[k] = bload(r_fw, G.cofs(x, j0 - (R - 1) + w + TW * k));
#pragma unroll
        for (int k = 0; k < NHF; ++k) {
            const int idx = min(tid + TT * k, HF * HCF - 1);
            const int r = idx / HCF, h = idx - r * HCF;
            fh[k] = bload(r_fw, G.cofs(i0 - (R - 1) + (h < R - 1 ? h : h + kTX), j0 - (R - 1) + r));
        }
    }
    // LDS stores
#pragma unroll
    for (int k = 0; k < NRU; ++k) {
        const int r = w + TW * k;
        if (r < HU) uu[r * WU + lane + H] = cu[k];
    }
#pragma unroll
    for (int k = 0; k < NHU; ++k) {
        const int idx = tid + TT * k;
        const int r = idx / (2 * H), h = idx - r * 2 * H;
        uu[idx < TZ * 2 * H ? (r + H) * WU + (h < H ? h : h + kTX) : HU * WU] = hu[k];
    }
    __syncthreads();

    // ---------------- phase 2: sigma from the FP64 differences of u
    auto frame = [](int v, int lo, int hi) { return (v >= lo && v <= hi) ? 1.f : 0.f; };
    // dx_v / dz_v of u at region cell `o` of uu
    auto dux = [&](int o) {
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) s += a.cx[m - 1] * (float)(uu[o + m] - uu[o + 1 - m]);
        return s;
    };
    auto duz = [&](int o) {
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) s += a.cz[m - 1] * (float)(uu[o + m * WU] - uu[o + (1 - m) * WU]);
        return s;
    };
    const float fxs = frame(x, R - 1, nx - R - 1);
#pragma unroll
    for (int k = 0; k < NRZ; ++k) {
        const int r = w + TW * k;  // sigma_zy row: z = j0 - R + r
        if (r >= HZ) break;
        const int z = j0 - R + r;
        szy[r * kTX + lane] = mz[k] * (duz((r - R + H) * WU + lane + H) * frame(z, R - 1, nz - R - 1));
        if (r >= R && r < R + TZ)  // a tile row: sigma_xy at the lane's own column
            sxy[(r - R) * WX + lane + R] = mz[k] * (dux((r - R + H) * WU + lane + H) * fxs);
    }
#pragma unroll
    for (int k = 0; k < NHM; ++k) {
        const int idx = min(tid + TT * k, TZ * 2 * R - 1);
        const int r = idx / (2 * R), h = idx - r * 2 * R;
        const int c = h < R ? h : h + kTX;  // sigma_xy column: x = i0 - R + c
        const float v = hm[k] * (dux((r + H) * WU + c - R + H) * frame(i0 - R + c, R - 1, nx - R - 1));
        sxy[tid + TT * k < TZ * 2 * R ? r * WX + c : TZ * WX] = v;
    }
    __syncthreads();

    // ---------------- phase 3: velocity and displacement of the tile's own nodes
    const int e0 = a.ps_off ? gld(a.ps_off, t) : 0, e1 = a.ps_off ? gld(a.ps_off, t + 1) : 0;
    const int f0 = a.rec_off ? gld(a.rec_off, t) : 0, f1 = a.rec_off ? gld(a.rec_off, t + 1) : 0;
    const float fxl = frame(x, R, nx - R - 1);
    const float axl = gld(a.axp, x + kGhostX);
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const int lz = w + TW * k, z = j0 + lz;
        const float* row = sxy + lz * WX + lane + R;
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) s += a.cx[m - 1] * (row[m - 1] - row[-m]);
        const float* col = szy + (lz + R) * kTX + lane;
        float q = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) q += a.cz[m - 1] * (col[(m - 1) * kTX] - col[-m * kTX]);
        float ds = s * fxl + q * frame(z, R, nz - R - 1);
        if (e1 > e0) {
            const int key = node_key(x, z);
            for (int e = e0; e < e1; ++e) {
                const int2 en = gld(a.ps_ent, e);
                if (en.x == key) ds += st.ps_amp ? gld(st.ps_amp, en.y) : st.ps_const;
            }
        }
        ds += msy[k];
        const float vn = (mv[k] + mb[k] * ds) * (axl * gld(a.azp, z + kGhost));
        const double uo = uu[(lz + H) * WU + lane + H];
        const double un = uo + (double)vn * a.dt64;
        G.put(a.o_v_out, x, z, vn);
        bstore8(G.rf, pofs8(a.o_u_out, pitch, x, z), un);
        const unsigned co = G.cofs(x, z);
        if (st.snap) bstore(rsrc(st.snap, cb), co, (float)uo);
        if (st.dsamp) bstore(rsrc(st.dsamp, cb), co, (float)un);
        if (f1 > f0 && st.seis_col) {
            const int key = node_key(x, z);
            for (int e = f0; e < f1; ++e) {
                const int2 en = gld(a.rec_ent, e);
                if (en.x == key) gst(st.seis_col, en.y, un);
            }
        }
    }

    // ---------------- phase 4: the previous step's structure-kernel product
    if (kp) {
        __syncthreads();  // phase 3 has read sigma: the snapshot takes its space
#pragma unroll
        for (int k = 0; k < NRF; ++k) {
            const int r = w + TW * k;
            if (r < HF) ff[r * WF + lane + R - 1] = fm[k];
        }
#pragma unroll
        for (int k = 0; k < NHF; ++k) {
            const int idx = tid + TT * k;
            const int r = idx / HCF, h = idx - r * HCF;
            ff[idx < HF * HCF ? r * WF + (h < R - 1 ? h : h + kTX) : HF * WF] = fh[k];
        }
        __syncthreads();
        const bool xin = x >= R - 1 && x <= nx - R - 1;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int lz = w + TW * k, z = j0 + lz;
            const unsigned co = G.cofs(x, z);
            const bool zin = z >= R - 1 && z <= nz - R - 1;
            if (co == kOutside || (!xin && !zin)) continue;
            const int ou = (lz + H) * WU + lane + H;
            const float* frow = ff + (lz + R - 1) * WF + lane + R - 1;
            double acc = 0.0;
            if (xin) {
                float sf = 0.f;
#pragma unroll
                for (int m = 1; m <= R; ++m) sf += a.cx[m - 1] * (frow[m] - frow[1 - m]);
                acc += (double)dux(ou) * (double)sf;
            }
            if (zin) {
                float sf = 0.f;
#pragma unroll
                for (int m = 1; m <= R; ++m) sf += a.cz[m - 1] * (frow[m * WF] - frow[(1 - m) * WF]);
                acc += (double)duz(ou) * (double)sf;
            }
            gst(a.kacc, (int)(co / 4u), kold[k] - acc * a.kscale);
        }
    }
}

template <int R, int MODE>
__global__ __launch_bounds__(kTileThreads) void k_utile(StepArgs a)
{
    const int g = blockIdx.x;
    const JobStep st{a.ps_amp, a.syn, a.snap, a.dsamp, a.seis_col, nullptr, nullptr, nullptr, a.kfwd,
                     a.ps_const, 0, 0};
    utile_step<R, MODE>(a, st, (g % kXCD) * a.tiles_per_xcd + g / kXCD);
}

template <int R, int MODE>
__global__ __launch_bounds__(kTileThreads) void k_utile_group(GroupArgs G)
{
    const int g = blockIdx.x;
    const int T = (g % kXCD) * G.tiles_per_xcd + g / kXCD;
    if (T >= G.total_tiles) return;
    int j = 0;
    while (j + 1 < G.n_jobs && G.tile_start[j + 1] <= T) ++j;  // wave-uniform scan
    j = __builtin_amdgcn_readfirstlane(j);
    const JobStep st = G.step[j];
    const StepArgs a = G.jobs[2 * st.slot + st.parity];
    utile_step<R, MODE>(a, st, T - G.tile_start[j]);
}

// The structure-kernel product of the last step of a displacement-form run (the others are fused
// into the next step): K_mu -= scale * (D u . D fwd) with u the FP64 displacement (padded), the sums
// formed as in utile_step.
template <int R>
__global__ __launch_bounds__(256) void k_ukmu(StepArgs a, const double* __restrict__ u, const float* __restrict__ fwd,
                                              double* __restrict__ kmu, double scale)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)a.nx * a.nz) return;
    const int x = (int)(c % a.nx), z = (int)(c / a.nx);
    const bool xin = x >= R - 1 && x <= a.nx - R - 1, zin = z >= R - 1 && z <= a.nz - R - 1;
    if (!xin && !zin) return;
    const long p = (long)z * a.pitch + x, pz = a.pitch, cz = a.nx;
    double acc = 0.0;
    if (xin) {
        float su = 0.f, sf = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) {
            su += a.cx[m - 1] * (float)(u[p + m] - u[p + 1 - m]);
            sf += a.cx[m - 1] * (fwd[c + m] - fwd[c + 1 - m]);
        }
        acc += (double)su * (double)sf;
    }
    if (zin) {
        float su = 0.f, sf = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) {
            su += a.cz[m - 1] * (float)(u[p + m * pz] - u[p + (1 - m) * pz]);
            sf += a.cz[m - 1] * (fwd[c + m * cz] - fwd[c + (1 - m) * cz]);
        }
        acc += (double)su * (double)sf;
    }
    kmu[c] -= acc * scale;
}

// K_mu -= fw_nth * (dx_v(u) .* dx_v(u_fwd) + dz_v(u) .* dz_v(u_fwd)) with u the new displacement
// (padded) and u_fwd the forward snapshot (compact); the products are formed in FP64.
// PF / PD: float / double pointer types (global-address-space pointers in the grouped launch)
template <int R, class PF, class PD>
__device__ __forceinline__ void kmu_tile(const StepArgs& a, PF u, PF fwd, PD kmu, double scale, int t)
{
    if (t >= a.n_tiles) return;
    const int x = (t % a.tiles_x) * kTX + (threadIdx.x & 63);
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    if (x >= a.nx) return;
    const bool xin = x >= R - 1 && x <= a.nx - R - 1;
    const int pz = a.pitch, cz = a.nx;
#pragma unroll 2
    for (int k = 0; k < kTileZ / 4; ++k) {
        const int z = (t / a.tiles_x) * kTileZ + w + 4 * k;
        if (z >= a.nz) break;
        const bool zin = z >= R - 1 && z <= a.nz - R - 1;
        if (!xin && !zin) continue;
        const int p = z * pz + x, c = z * cz + x;
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
}

// the step kernel's 64 x 32 tiles in the same XCD-banded order, so the z-stencil rows a tile shares
// with the tiles above and below come from the same L2
template <int R>
__global__ __launch_bounds__(256) void k_kmu(StepArgs a, const float* __restrict__ u, const float* __restrict__ fwd,
                                             double* __restrict__ kmu, double scale)
{
    const int g = blockIdx.x;
    kmu_tile<R>(a, u, fwd, kmu, scale, (g % kXCD) * a.tiles_per_xcd + g / kXCD);  // kernel arguments: global
}

void launch_kmu(int R, const StepArgs& a, const float* u, const float* fwd, double* kmu, double scale, hipStream_t s)
{
    const dim3 g((unsigned)(kXCD * a.tiles_per_xcd)), b(256);
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

void launch_pair_group(int R, bool with_u, const PairGroupArgs& ga, hipStream_t s)
{
    const dim3 g((unsigned)(kXCD * ga.tiles_per_xcd)), b(kPairThreads);
#define FD2D_PGROUP(RR)                                                                          \
    case RR:                                                                                     \
        if (with_u) hipLaunchKernelGGL((k_pair_group<RR, 1>), g, b, 0, s, ga);                   \
        else hipLaunchKernelGGL((k_pair_group<RR, 0>), g, b, 0, s, ga);                          \
        break;
    switch (R) {
        FD2D_PGROUP(1)
        FD2D_PGROUP(2)
        FD2D_PGROUP(3)
        FD2D_PGROUP(4)
    }
#undef FD2D_PGROUP
}

void launch_pair(int R, bool with_u, const PairArgs& p, hipStream_t s)
{
    const dim3 g((unsigned)(kXCD * p.a.tiles_per_xcd)), b(kPairThreads);
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

void launch_group(int R, bool with_u, const GroupArgs& ga, hipStream_t s)
{
    const dim3 g((unsigned)(kXCD * ga.tiles_per_xcd)), b(kTileThreads);
    bool product = false;  // some solve of the group takes a structure-kernel product this step
    for (int j = 0; j < ga.n_jobs; ++j) product = product || ga.step[j].kfwd != nullptr;
#define FD2D_GROUP(RR)                                                                           \
    case RR:                                                                                     \
        if (with_u && product) hipLaunchKernelGGL((k_tile_group<RR, 2>), g, b, 0, s, ga);        \
        else if (with_u) hipLaunchKernelGGL((k_tile_group<RR, 1>), g, b, 0, s, ga);              \
        else hipLaunchKernelGGL((k_tile_group<RR, 0>), g, b, 0, s, ga);                          \
        break;
    switch (R) {
        FD2D_GROUP(1)
        FD2D_GROUP(2)
        FD2D_GROUP(3)
        FD2D_GROUP(4)
    }
#undef FD2D_GROUP
}

void launch_step(int R, bool with_u, const StepArgs& a, hipStream_t s)
{
    const dim3 g((unsigned)(kXCD * a.tiles_per_xcd)), b(kTileThreads);
#define FD2D_TILE(RR)                                                                            \
    case RR:                                                                                     \
        if (with_u && a.kfwd) hipLaunchKernelGGL((k_tile<RR, 2>), g, b, 0, s, a);                \
        else if (with_u) hipLaunchKernelGGL((k_tile<RR, 1>), g, b, 0, s, a);                     \
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

void launch_ustep(int R, const StepArgs& a, hipStream_t s)
{
    const dim3 g((unsigned)(kXCD * a.tiles_per_xcd)), b(kTileThreads);
    if (R == 3) {
        if (a.kfwd) hipLaunchKernelGGL((k_utile<3, 1>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_utile<3, 0>), g, b, 0, s, a);
    } else {
        if (a.kfwd) hipLaunchKernelGGL((k_utile<4, 1>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_utile<4, 0>), g, b, 0, s, a);
    }
}

void launch_ugroup(int R, const GroupArgs& ga, hipStream_t s)
{
    const dim3 g((unsigned)(kXCD * ga.tiles_per_xcd)), b(kTileThreads);
    bool product = false;
    for (int j = 0; j < ga.n_jobs; ++j) product = product || ga.step[j].kfwd != nullptr;
    if (R == 3) {
        if (product) hipLaunchKernelGGL((k_utile_group<3, 1>), g, b, 0, s, ga);
        else hipLaunchKernelGGL((k_utile_group<3, 0>), g, b, 0, s, ga);
    } else {
        if (product) hipLaunchKernelGGL((k_utile_group<4, 1>), g, b, 0, s, ga);
        else hipLaunchKernelGGL((k_utile_group<4, 0>), g, b, 0, s, ga);
    }
}

void launch_ukmu(int R, const StepArgs& a, const double* u, const float* fwd, double* kmu, double scale,
                 hipStream_t s)
{
    const long n = (long)a.nx * a.nz;
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
    if (R == 3) hipLaunchKernelGGL(k_ukmu<3>, g, b, 0, s, a, u, fwd, kmu, scale);
    else hipLaunchKernelGGL(k_ukmu<4>, g, b, 0, s, a, u, fwd, kmu, scale);
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
// binfo: {dt/rho uniform (1/0), its value, mu uniform, its value, (int) dt/rho differs, (int) mu differs}
__global__ void k_binfo_init(float* binfo)
{
    int* d = (int*)(binfo + 4);
    d[0] = 0;
    d[1] = 0;
}

__global__ void k_material(const double* mu, const double* rho, float* bv, float* dtmu, int nx, int nz,
                           int pitch, double dt, double mu_scale, float* binfo)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)nx * nz) return;
    const int i = (int)(c % nx), j = (int)(c / nx);
    const long p = (long)j * pitch + i;
    const float b = (float)(dt / rho[c]), m = (float)(mu_scale * mu[c]);
    bv[p] = b;
    dtmu[p] = m;
    // a value other than node 0's marks the coefficient as varying (one store per wavefront)
    const bool db = !(b == (float)(dt / rho[0])), dm = !(m == (float)(mu_scale * mu[0]));
    int* d = (int*)(binfo + 4);
    if (__any(db) && (threadIdx.x & 63) == __builtin_ffsll(__ballot(1)) - 1) __hip_atomic_store(d, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (__any(dm) && (threadIdx.x & 63) == __builtin_ffsll(__ballot(1)) - 1) __hip_atomic_store(d + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_binfo_final(const double* mu, const double* rho, double dt, double mu_scale, float* binfo)
{
    const int* d = (const int*)(binfo + 4);
    binfo[0] = d[0] ? 0.f : 1.f;
    binfo[1] = (float)(dt / rho[0]);
    binfo[2] = d[1] ? 0.f : 1.f;
    binfo[3] = (float)(mu_scale * mu[0]);
}

void launch_material(const double* mu, const double* rho, float* bv, float* dtmu, int nx, int nz,
                     int pitch, double dt, double mu_scale, float* binfo, hipStream_t s)
{
    const long n = (long)nx * nz;
    hipLaunchKernelGGL(k_binfo_init, dim3(1), dim3(1), 0, s, binfo);
    hipLaunchKernelGGL(k_material, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mu, rho, bv, dtmu,
                       nx, nz, pitch, dt, mu_scale, binfo);
    hipLaunchKernelGGL(k_binfo_final, dim3(1), dim3(1), 0, s, mu, rho, dt, mu_scale, binfo);
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

// Finiteness check of a solve's field pool: partial[b] = 1 when block b saw a NaN or Inf among its
// elements (grid-stride over float4; n4 = elements / 4), then flag = OR of the partials. Every
// partial and the flag are written on every call, so no clearing is needed between checks.
__global__ __launch_bounds__(256) void k_nonfinite_partial(const float4* __restrict__ x, size_t n4,
                                                           int* __restrict__ partial)
{
    int bad = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)kCheckBlocks * 256) {
        const float4 q = x[i];
        // a comparison with NaN is false; Inf exceeds the largest finite float
        bad |= (int)!(fabsf(q.x) <= FLT_MAX) | (int)!(fabsf(q.y) <= FLT_MAX) | (int)!(fabsf(q.z) <= FLT_MAX) |
               (int)!(fabsf(q.w) <= FLT_MAX);
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) partial[blockIdx.x] = bad ? 1 : 0;
}

__global__ __launch_bounds__(kCheckBlocks) void k_nonfinite_final(const int* __restrict__ partial,
                                                                  int* __restrict__ flag)
{
    const int bad = __syncthreads_or(partial[threadIdx.x]);
    if (threadIdx.x == 0) *flag = bad ? 1 : 0;
}

void launch_nonfinite(const float* x, size_t n, int* partial, int* flag, hipStream_t s)
{
    hipLaunchKernelGGL(k_nonfinite_partial, dim3(kCheckBlocks), dim3(256), 0, s, (const float4*)x, n / 4, partial);
    hipLaunchKernelGGL(k_nonfinite_final, dim3(1), dim3(kCheckBlocks), 0, s, (const int*)partial, flag);
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
// a[2k][x], a[2k+1][x] = Re, Im of A[k][x] = spec_sum[k] * dist(x,k) * (conj_in ? conj(Z) : Z)[k][x],
// dist(x,k) = sum_b I[k,b] sd[b][x], formed in FP64 and rounded to FP32 as the synthesis reads them;
// rows 2K..C_pad-1 are zero.
// One thread per node and frequency block (kSpecKB frequencies of the padded C_pad / 2): the rows of
// a node are written by C_pad / (2 kSpecKB) workgroups side by side, so every store is coalesced
// and a small grid still fills the chip.
constexpr int kSpecKB = 8;
__global__ void k_source_spectrum(float* a32, const double2* Z, int conj_in, const double* sd, int n_sd,
                                  const double* I, const double* spec_sum, size_t P, int K, int C_pad)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= P) return;
    const int k0 = blockIdx.y * kSpecKB;
#pragma unroll
    for (int q = 0; q < kSpecKB; ++q) {
        const int k = k0 + q;
        if (2 * k >= C_pad) break;
        if (k >= K) {  // zero rows 2K .. C_pad - 1
            a32[(size_t)(2 * k) * P + x] = 0.f;
            a32[(size_t)(2 * k + 1) * P + x] = 0.f;
            continue;
        }
        double d = 0.0;
        for (int b = 0; b < n_sd; ++b) d += I[k + (size_t)K * b] * sd[x + P * b];
        const double a = spec_sum[k] * d;
        const double2 z = Z[(size_t)k * P + x];
        a32[(size_t)(2 * k) * P + x] = (float)(a * z.x);
        a32[(size_t)(2 * k + 1) * P + x] = (float)(conj_in ? -(a * z.y) : a * z.y);
    }
}

void launch_source_spectrum(float* a32, const double2* Z, int conj_in, const double* sd, int n_sd,
                            const double* I, const double* spec_sum, size_t P, int K, int C_pad, hipStream_t s)
{
    const unsigned nkb = (unsigned)((C_pad / 2 + kSpecKB - 1) / kSpecKB);
    hipLaunchKernelGGL(k_source_spectrum, dim3((unsigned)((P + 255) / 256), nkb), dim3(256), 0, s, a32, Z, conj_in,
                       sd, n_sd, I, spec_sum, P, K, C_pad);
}

__global__ void k_finalize_dft(double2* out, const double2* acc, const double2* comp, size_t n, double scale,
                               int conj)
{
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double2 a = acc[c], e = comp ? comp[c] : make_double2(0.0, 0.0);
    const double re = a.x + e.x, im = a.y + e.y;
    out[c] = make_double2(scale * re, conj ? -(scale * im) : scale * im);
}

void launch_finalize_dft(double2* out, const double2* acc, const double2* comp, size_t n, double scale, int conj,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_finalize_dft, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, acc, comp, n, scale,
                       conj);
}

__global__ void k_source_kernel(double* Ks, const double2* G, const double2* Dacc, const double2* Dcomp,
                                const double* I, const double* spec, size_t P, int K, int nb_cols, double scale)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= P) return;
    for (int b = 0; b < nb_cols; ++b) {
        double s = 0.0;
        for (int k = 0; k < K; ++k) {
            const double2 g = G[(size_t)k * P + x], da = Dacc[(size_t)k * P + x];
            const double2 dc = Dcomp ? Dcomp[(size_t)k * P + x] : make_double2(0.0, 0.0);
            const double2 d = make_double2(da.x + dc.x, da.y + dc.y);
            s += I[k + (size_t)K * b] * spec[k] * (g.x * d.x - g.y * d.y);
        }
        Ks[(size_t)b * P + x] = scale * s;
    }
}

void launch_source_kernel(double* Ks, const double2* G, const double2* Dacc, const double2* Dcomp, const double* I,
                          const double* spec, size_t P, int K, int nb_cols, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_source_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, Ks, G, Dacc, Dcomp, I,
                       spec, P, K, nb_cols, scale);
}

}  // namespace fd2d
