// engine.h -- device-side data layout and kernel launch interface of the gfx950 SH engine.
//
// Wavefield layout in HBM: every time-stepped field (v, sigma_xy, sigma_zy, u) and every per-node
// coefficient (dt/rho, dt*mu) is an FP32 nx x nz array, x fastest (the MATLAB column-major order,
// so a wavefront reads 64 consecutive x nodes of one z row: 256 contiguous bytes). Each array is
// padded with kGhost zero rows above and below, kGhostX zero columns on the left and a row pitch
// rounded up to 64 elements plus 64, so every load of a tile's halo has a non-negative byte offset
// and reads zeros outside the grid (loads past the end are range-checked buffer loads). sigma_xy and sigma_zy share the
// node (i,j) of the MATLAB arrays sxy(i,j) ((nx-1) x nz) and szy(i,j) (nx x (nz-1)).
//
// The time-stepped fields are double-buffered: step n reads the "in" set and writes the "out" set,
// so a workgroup never reads a neighbour's already-updated values.
//
// Snapshots (G_out, C_out, u_out, u_fwd, G_in, C_in, u_in), DFT samples and synthesised source
// slices are compact nx x nz arrays (the MATLAB layout, so they are copied out unchanged).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fd2d {

constexpr int kGhost = 16;  // ghost rows (>= 4R-2 for R <= 4: the sigma halo of a two-step launch)
constexpr int kGhostX = 32; // ghost columns: node x = 0 starts a 128-byte line
constexpr int kTX = 64;     // tile width: one wavefront per row
constexpr int kXCD = 8;     // MI355X accelerator complex dies (workgroups are dealt round-robin)
constexpr int kTileZ = 32;  // rows of a step tile
constexpr int kTileThreads = 512;  // threads of a step workgroup (8 wavefronts, rows w, w+8, ...)
constexpr int kPairThreads = 512;  // threads of a two-step workgroup (8 wavefronts, rows w, w+8, ...)
constexpr int kMaxDim = 32000;  // node keys pack (x+8, z+8) into 16 bits each

struct PaddedLayout {
    int nx = 0, nz = 0, pitch = 0, rows = 0;
    size_t elems() const { return (size_t)pitch * (size_t)rows; }
    size_t origin() const { return (size_t)kGhost * (size_t)pitch + kGhostX; }
    static PaddedLayout make(int nx, int nz)
    {
        PaddedLayout l;
        l.nx = nx;
        l.nz = nz;
        // whole tiles plus every halo of a one- or two-step launch (<= 4R-2 <= kGhost rows and
        // columns) stay inside the array: kGhostX columns on the left, >= 32 on the right of the
        // last tile, kGhost rows above and below the last tile row
        const int tiles_x = (nx + kTX - 1) / kTX, tiles_z = (nz + kTileZ - 1) / kTileZ;
        l.pitch = kTX * (tiles_x + 1);
        l.rows = kGhost + kTileZ * tiles_z + kGhost;
        return l;
    }
};

// Arguments of one fused time step.
struct StepArgs {
    // state (origin pointers into padded arrays); in = step n-1, out = step n
    const float* v_in;
    const float* sxy_in;
    const float* szy_in;
    const float* u_in;   // displacement (correlation / adjoint runs), else null
    float* v_out;
    float* sxy_out;
    float* szy_out;
    float* u_out;
    const float* bv;     // dt / rho
    const float* dtmu;   // dt * mu
    const float* binfo;  // {dt/rho uniform, its value, mu uniform, its value}: see make_grid
    const float* ax;     // separable taper, nx values
    const float* az;     // nz values
    const float* axp;    // the taper padded like the fields: axp[x + kGhostX], azp[z + kGhost], zero
    const float* azp;    // outside the grid
    // the padded arrays as buffer pools: fields (both ping-pong sets) and coefficients (dt/rho,
    // dt*mu); o_* are element offsets of each padded array (its row -kGhost, column -kGhostX)
    const float* fbase;
    unsigned fbytes;
    const float* cbase;
    unsigned cbytes;
    int o_v_in, o_sxy_in, o_szy_in, o_u_in, o_v_out, o_sxy_out, o_szy_out, o_u_out, o_bv, o_dtmu;
    unsigned compact_bytes;  // nx*nz*4: the range of the compact arrays
    int nx, nz, pitch;
    int tiles_x, n_tiles;           // tiles across x, tiles in all
    int tiles_per_xcd;              // tile t = (g % 8) * tiles_per_xcd + g / 8 for workgroup g
    int origin;                     // element offset of node (0,0) in a padded array
    unsigned padded_bytes;          // bytes of one padded array
    float cx[4], cz[4];  // staggered coefficients / dx, / dz
    float dt;

    // point sources: per-tile CSR over the nodes whose velocity the tile computes (tile plus
    // halo); entries {node key ((z+8)<<16 | (x+8)), source}
    const int* ps_off;
    const int2* ps_ent;
    const float* ps_amp;     // amplitudes of this step indexed by source (null: ps_const)
    float ps_const;
    const float* syn;        // synthesised distributed source of this step (compact), or null
    // With u present (correlation / adjoint) the snapshot is u before the update and the DFT
    // sample and receivers take u after it; without u (Green) they take v.
    float* snap;             // snapshot destination (compact) or null
    float* dsamp;            // DFT sample destination (compact) or null
    // receivers: per-tile CSR over its own nodes; entries {node key, receiver}
    const int* rec_off;
    const int2* rec_ent;
    double* seis_col;        // seismograms(:, n)

    // stress-phase Born terms
    const float* born_dmu;   // dmu as FP32 (compact) for C_in / u_in Born terms, or null
    const float* born_in0;   // C_in(:,:,n)
    const float* born_in1;   // C_in(:,:,n+1)

    // structure-kernel product fused into the step (adjoint runs): K_mu -= kscale * (D u . D kfwd)
    const float* kfwd;       // forward snapshot of this step (compact), or null
    double* kacc;            // K_mu accumulator (compact, FP64)
    double kscale;           // fw_nth
    double dt64;             // dt in FP64 (u_n = u_{n-1} + dt v_n of the displacement-form step)
};

inline int node_key_host(int x, int z) { return ((z + 8) << 16) | (x + 8); }

// Two consecutive time steps n, n+1 in one launch. `a` describes the state before step n, the
// outputs after step n+1 and the per-step inputs of step n (ps_amp / ps_const, syn, snap, dsamp,
// seis_col); the fields below are those of step n+1. At most one of the two steps may carry a
// synthesised source (a.syn or syn2); Born terms are not supported.
struct PairArgs {
    StepArgs a;
    const float* ps_amp2;
    float ps_const2;
    const float* syn2;
    float* snap2;
    float* dsamp2;
    double* seis_col2;
};

// The per-step fields of one solve in a grouped launch (the rest of its StepArgs is fixed).
struct JobStep {
    const float* ps_amp;
    const float* syn;
    float* snap;
    float* dsamp;
    double* seis_col;
    const float* born_dmu;  // stress-phase Born terms of this step (null: none)
    const float* born_in0;
    const float* born_in1;
    const float* kfwd;      // forward snapshot of this step's structure-kernel product (null: none)
    float ps_const;
    int parity;  // which of the solve's two StepArgs (ping-pong direction) this step uses
    int slot;    // the solve's StepArgs are jobs[2 * slot + parity] of the grouped launch
};

constexpr int kGroupMax = 40;  // solves per grouped launch (kernel arguments stay under 4 KB)

// One time step of up to kGroupMax solves with the same order and field set in one launch: the
// tiles of all solves are concatenated (solve j owns tiles tile_start[j] .. tile_start[j+1]-1)
// and dealt to the XCDs in contiguous bands. Solve j's StepArgs are jobs[2 * step[j].slot + parity].
struct GroupArgs {
    const StepArgs* jobs;
    int n_jobs;
    int total_tiles;
    int tiles_per_xcd;
    int tile_start[kGroupMax + 1];
    JobStep step[kGroupMax];
};
static_assert(sizeof(GroupArgs) <= 4096, "grouped-launch arguments exceed the 4 KB kernel-argument limit");

// The per-step fields of one solve in a grouped two-step launch: those of steps n and n+1.
struct PairStep {
    const float* ps_amp;
    const float* syn;
    float* snap;
    float* dsamp;
    double* seis_col;
    const float* ps_amp2;
    const float* syn2;
    float* snap2;
    float* dsamp2;
    double* seis_col2;
    float ps_const;
    float ps_const2;
    int parity;
    int slot;
};
constexpr int kPairGroupMax = 36;  // solves per grouped two-step launch (arguments under 4 KB)
struct PairGroupArgs {
    const StepArgs* jobs;
    int n_jobs;
    int total_tiles;
    int tiles_per_xcd;
    int tile_start[kPairGroupMax + 1];
    PairStep step[kPairGroupMax];
};
static_assert(sizeof(PairGroupArgs) <= 4096, "grouped-launch arguments exceed the 4 KB kernel-argument limit");
void launch_pair_group(int R, bool with_u, const PairGroupArgs& g, hipStream_t s);

// halo of the first step's velocity region of a pair launch (the point-source lists cover it)
inline int pair_halo_lo(int R) { return 3 * R - 2; }
inline int pair_halo_hi(int R) { return 3 * R - 1; }

void launch_step(int R, bool with_u, const StepArgs& a, hipStream_t s);
void launch_pair(int R, bool with_u, const PairArgs& p, hipStream_t s);
void launch_group(int R, bool with_u, const GroupArgs& g, hipStream_t s);

// K_mu -= scale * (D u . D fwd) over the grid (u padded, fwd compact), after a step.
void launch_kmu(int R, const StepArgs& a, const float* u, const float* fwd, double* kmu, double scale,
                hipStream_t s);
// Green-function Born source applied at the start of a step: sigma += (dt*dmu) .* D(G_in slice),
// in place on the current state.
void launch_born_green(int R, const StepArgs& a, float* sxy, float* szy, const float* dtdmu, const float* gin,
                       hipStream_t s);

// Material coefficients: bv = dt/rho, dtmu = mu_scale*mu (dt*mu for the velocity-stress step, mu for the
// displacement-form step) into padded arrays, and binfo = {dt/rho uniform, its value, mu uniform, its
// value, two int flags} (6 words).
void launch_material(const double* mu, const double* rho, float* bv, float* dtmu, int nx, int nz,
                     int pitch, double dt, double mu_scale, float* binfo, hipStream_t s);

// Displacement-form steps (correlation and adjoint runs without Born terms, R = 3 or 4): the field
// pool holds v (FP32) and u (FP64) only, o_u_in / o_u_out are FP32-element offsets of the FP64
// arrays, o_dtmu holds mu. a.kfwd (or step[j].kfwd) is the forward snapshot of the PREVIOUS step's
// structure-kernel product, which these launches compute from their staged input displacement.
void launch_ustep(int R, const StepArgs& a, hipStream_t s);
void launch_ugroup(int R, const GroupArgs& g, hipStream_t s);
// the product of the last step of a displacement-form run, from its FP64 displacement u (origin pointer)
void launch_ukmu(int R, const StepArgs& a, const double* u, const float* fwd, double* kmu, double scale,
                 hipStream_t s);
void launch_scale_to_float(const double* in, float* out, size_t n, double scale, hipStream_t s);

// acc[k][x] + comp[k][x] += sum_j samp[j][x] * e[j][k] over cnt samples: matrix-core products of
// the three bfloat16 parts of each FP32 operand (FP32-accurate) summed in FP32 over the slab, then
// added into the FP64 accumulators with Neumaier compensation (comp holds the compensation terms;
// the sum is acc + comp); e{h,m,l}[c][Jt] are the parts of Re e[j][k] (c = 2k) and Im e[j][k]
// (c = 2k + 1), C_pad (a multiple of 256) rows, offset to the slab's first sample (16-byte aligned).
// first != 0: the sums start from zero (acc and comp are not read). out != nullptr: the last slab
// of the series; instead of acc and comp, out[k][x] = scale * (conj ? conj(sum) : sum) is written
// (out may be acc; comp is not written and may be null when first != 0).
void launch_dft_bf16(double2* acc, double2* comp, const float* samp, const unsigned short* eh,
                     const unsigned short* em, const unsigned short* el, size_t P, int K, int C_pad, int Jt,
                     int cnt, int first, double2* out, double scale, int conj, hipStream_t s);
// out[j][x] = sum_c et[j][c] * a[c][x] for j < cnt, the same products: with a = (Re A, Im A) rows
// and et = (Re, -Im) of the phase factors (parts t{h,m,l}[J_pad][C_pad], J_pad >= cnt rounded up to
// 64) this is Re(A[k][x] * coef[j][k]) summed over k.
void launch_synth_bf16(float* out, const float* a, const unsigned short* th, const unsigned short* tm,
                       const unsigned short* tl, size_t P, int C_pad, int cnt, hipStream_t s);
// the three bfloat16 parts (rounded to nearest, v = h + m + l exactly) of an FP32 value
void split_bf16(float v, unsigned short& h, unsigned short& m, unsigned short& l);
// a32[2k][x], a32[2k+1][x] = Re, Im of A[k][x] = spec_sum[k] * dist(x,k) * (conj_in ? conj(Z) : Z)[k][x]
// (FP64, rounded to FP32), dist(x,k) = sum_b I[k,b] sd[b][x]; zero rows 2K..C_pad-1
void launch_source_spectrum(float* a32, const double2* Z, int conj_in, const double* sd, int n_sd,
                            const double* I, const double* spec_sum, size_t P, int K, int C_pad, hipStream_t s);
// out[k][x] = scale * (conj ? conj(acc + comp) : acc + comp)  (out may be acc; comp may be null: 0)
void launch_finalize_dft(double2* out, const double2* acc, const double2* comp, size_t n, double scale, int conj,
                         hipStream_t s);
// Ks[b][x] = scale * sum_k I[k,b] spec[k] Re(conj(G[k][x] * D[k][x])), D = Dacc + Dcomp (Dcomp may be
// null: 0)
void launch_source_kernel(double* Ks, const double2* G, const double2* Dacc, const double2* Dcomp, const double* I,
                          const double* spec, size_t P, int K, int nb_cols, double scale,
                          hipStream_t s);
// *flag = 1 if any of the n floats at x (n a multiple of 4, x 16-byte aligned) is NaN or Inf,
// else 0; partial holds kCheckBlocks ints of scratch
constexpr int kCheckBlocks = 256;
void launch_nonfinite(const float* x, size_t n, int* partial, int* flag, hipStream_t s);
// point-source amplitudes: amp[n][q] = (float)Re(stf(q, n)) from complex nr x N interleaved
void launch_stf_real(float* amp, const double2* stf, int64_t d0, int nq, int nsteps, hipStream_t s);

constexpr int kSpecSlab = 1024;  // time samples per DFT flush / synthesis block (at most)

}  // namespace fd2d
