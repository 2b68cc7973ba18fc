// capi.hip -- C ABI of include/fd2d_noise.h: argument checks with the reference's empty-array
// semantics, device set-up, and the time loops of the three solvers on one HIP stream.
//
// The loops follow the step order of code/run_forward1_green.m:130-192,
// code/run_forward2_correlation.m:178-251 and code/run_noise_adjoint.m:233-355; every kernel they
// launch is in kernels.hip.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fd2d_noise.h"
#include "caller.h"
#include "engine.h"

using namespace fd2d;

namespace {

thread_local std::string g_err;

struct Failure {
    int code;
};

[[noreturn]] void raise(int code, const char* fmt, ...)
{
    char buf[768];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    throw Failure{code};
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) raise(FD2D_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_));     \
    } while (0)

// Stream of the work the current entry point enqueues (StreamScope); device buffers remember it so
// a released block is known free once that stream has passed the release point.
thread_local hipStream_t t_stream = nullptr;
struct StreamScope {
    hipStream_t prev;
    explicit StreamScope(hipStream_t s) : prev(t_stream) { t_stream = s; }
    ~StreamScope() { t_stream = prev; }
    StreamScope(const StreamScope&) = delete;
    StreamScope& operator=(const StreamScope&) = delete;
};

// Process-wide cache of released device blocks, per device. hipMalloc / hipFree of multi-GB
// buffers cost tens of milliseconds each on a device whose memory other processes used before
// (and hipFree waits for the whole device), and every solve allocates the same few sizes again, so
// released blocks are kept and handed out again for requests of at least 2/3 of their size. A
// released block carries an event recorded on the stream that used it; the next user waits on the
// event, so a block still read by queued work is never handed out early. At most `limit` bytes are
// kept (a quarter of the device memory), the blocks released longest ago freed first to make room;
// fd2d_release_cached_memory() frees them all. A stream
// that is about to be destroyed is detached first: the events recorded on it are waited for and
// destroyed, so that no cached event refers to a stream that no longer exists.
struct BlockCache {
    struct Block {
        void* p;
        hipEvent_t ev;
        hipStream_t st;  // the stream ev was recorded on
        uint64_t seq;    // order of release: the cache evicts the least recently released first
    };
    std::mutex mu;
    std::multimap<size_t, Block> blocks;
    size_t held = 0, limit = 0;
    uint64_t next_seq = 0;
    bool limit_set = false;
    void* take(size_t bytes, size_t* got)
    {
        Block blk{nullptr, nullptr, nullptr};
        {
            std::lock_guard<std::mutex> lk(mu);
            auto it = blocks.lower_bound(bytes);
            if (it == blocks.end() || it->first > bytes + bytes / 2) return nullptr;
            blk = it->second;
            *got = it->first;
            held -= it->first;
            blocks.erase(it);
        }
        if (blk.ev) {
            if (hipEventSynchronize(blk.ev) != hipSuccess) (void)hipGetLastError();
            (void)hipEventDestroy(blk.ev);
        }
        return blk.p;
    }
    void give(void* p, size_t bytes, hipStream_t st)
    {
        std::unique_lock<std::mutex> lk(mu);
        if (!limit_set) {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) limit = total_b / 4;
            limit_set = true;
        }
        if (bytes > limit) {
            lk.unlock();
            (void)hipFree(p);
            return;
        }
        // Full: make room by freeing the blocks released longest ago, so that the cache follows the
        // sizes in use now. (Keeping the old blocks and freeing each new one instead held a cache
        // full of one workload's sizes while the next workload paid a hipMalloc and a hipFree, which
        // waits for the device, for every block.)
        while (held + bytes > limit) {
            auto old = blocks.begin();
            for (auto it = blocks.begin(); it != blocks.end(); ++it)
                if (it->second.seq < old->second.seq) old = it;
            if (old->second.ev) {
                if (hipEventSynchronize(old->second.ev) != hipSuccess) (void)hipGetLastError();
                (void)hipEventDestroy(old->second.ev);
            }
            (void)hipFree(old->second.p);
            held -= old->first;
            blocks.erase(old);
        }
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(ev, st) != hipSuccess) {
            (void)hipGetLastError();
            if (ev) (void)hipEventDestroy(ev);
            lk.unlock();
            if (hipStreamSynchronize(st) != hipSuccess) (void)hipGetLastError();
            (void)hipFree(p);
            return;
        }
        blocks.emplace(bytes, Block{p, ev, st, next_seq++});
        held += bytes;
    }
    void detach(hipStream_t s)
    {
        std::lock_guard<std::mutex> lk(mu);
        for (auto& kv : blocks) {
            Block& b = kv.second;
            if (b.ev && b.st == s) {
                if (hipEventSynchronize(b.ev) != hipSuccess) (void)hipGetLastError();
                (void)hipEventDestroy(b.ev);
                b.ev = nullptr;
                b.st = nullptr;
            }
        }
    }
    void clear()
    {
        std::lock_guard<std::mutex> lk(mu);
        for (auto& kv : blocks) {
            if (kv.second.ev) {
                if (hipEventSynchronize(kv.second.ev) != hipSuccess) (void)hipGetLastError();
                (void)hipEventDestroy(kv.second.ev);
            }
            (void)hipFree(kv.second.p);
        }
        blocks.clear();
        held = 0;
    }
};

constexpr int kMaxDevices = 64;
BlockCache g_cache[kMaxDevices];

BlockCache* device_cache()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return nullptr;
    }
    return &g_cache[dev];
}

// Owning device buffer: a block of the device's BlockCache, else hipMalloc. An upload keeps its host
// copy (`stage`) until the buffer is released.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0, bytes = 0;
    BlockCache* cache = nullptr;
    hipStream_t st = nullptr;
    std::vector<unsigned char> stage;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        bytes = count * sizeof(T);
        cache = device_cache();
        st = t_stream;
        if (cache) {
            size_t got = 0;
            if (void* q = cache->take(bytes, &got)) {
                p = (T*)q;
                bytes = got;
                n = count;
                return;
            }
        }
        if (hipMalloc((void**)&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            // memory held by the cache may be what is missing: return it and try once more
            if (cache) cache->clear();
            if (hipMalloc((void**)&p, bytes) != hipSuccess) {
                p = nullptr;
                (void)hipGetLastError();
                raise(FD2D_ENOMEM, "device allocation of %zu bytes failed", bytes);
            }
        }
        n = count;
    }
    void release()
    {
        if (p) {
            if (cache) cache->give(p, bytes, st);
            else (void)hipFree(p);
        }
        p = nullptr;
        n = bytes = 0;
        cache = nullptr;
        stage.clear();
        stage.shrink_to_fit();
    }
    ~DBuf() { release(); }
};

int64_t numel(const fd2d_carray* a)
{
    if (!a || !a->data) return 0;
    return a->dims[0] * a->dims[1] * a->dims[2];
}
int64_t numel(const fd2d_marray* a)
{
    if (!a || !a->data) return 0;
    return a->dims[0] * a->dims[1] * a->dims[2];
}
// MATLAB size(A,d) with [] = 0 x 0 (size 1 along the third dimension)
int64_t dim(const fd2d_carray* a, int d) { return numel(a) ? a->dims[d] : (d == 2 ? 1 : 0); }

// Staggered first-derivative coefficients. Orders 2 and 4 are the ones of
// code/propagation/dx_v.m:23,29; 6 and 8 are the Taylor coefficients of the same stencil family.
const double kCoef[5][4] = {{0, 0, 0, 0},
                            {1.0, 0, 0, 0},
                            {9.0 / 8.0, -1.0 / 24.0, 0, 0},
                            {75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0, 0},
                            {1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0}};

struct Geometry {
    int nx, nz, R, K, s, f, nt;
    double dx, dz, dt, dw;
    size_t P;
};

Geometry geometry(const fd2d_config* cfg)
{
    if (!cfg) raise(FD2D_EINVAL, "config is NULL");
    Geometry g{};
    g.nx = cfg->nx;
    g.nz = cfg->nz;
    if (g.nx < 3 || g.nz < 3) raise(FD2D_EINVAL, "nx and nz must be >= 3 (got %d x %d)", g.nx, g.nz);
    if (g.nx > kMaxDim || g.nz > kMaxDim) raise(FD2D_EUNSUPPORTED, "nx and nz must be <= %d", kMaxDim);
    if (cfg->order != 2 && cfg->order != 4 && cfg->order != 6 && cfg->order != 8)
        raise(FD2D_EUNSUPPORTED, "FD order %d is not one of 2, 4, 6, 8", cfg->order);
    g.R = cfg->order / 2;
    g.K = cfg->n_sample;
    g.s = cfg->freq_samp;
    g.f = cfg->fw_nth;
    g.nt = cfg->nt;
    if (g.s < 1 || g.f < 1 || g.nt < 1) raise(FD2D_EINVAL, "freq_samp, fw_nth and nt must be >= 1");
    if (g.K < 0 || (g.K > 0 && !cfg->w_sample)) raise(FD2D_EINVAL, "w_sample must hold n_sample values");
    if (!(cfg->Lx > 0) || !(cfg->Lz > 0) || !(cfg->dt > 0)) raise(FD2D_EINVAL, "Lx, Lz and dt must be positive");
    if (cfg->n_basis_fct < 0 || cfg->n_basis_fct == 1 || (cfg->n_basis_fct != 0 && cfg->n_basis_fct > g.K))
        raise(FD2D_EUNSUPPORTED, "n_basis_fct=%d: the band interpolation needs 0 or 2..n_sample", cfg->n_basis_fct);
    g.dx = cfg->Lx / (cfg->nx - 1);
    g.dz = cfg->Lz / (cfg->nz - 1);
    g.dt = cfg->dt;
    g.dw = cfg->dw;
    g.P = (size_t)g.nx * (size_t)g.nz;
    return g;
}

// Nearest grid node per axis, lowest index on ties (code/run_forward1_green.m:45-60).
int nearest_node(double p, double h, int n)
{
    int best = 0;
    double bd = std::fabs(0.0 - p);
    for (int i = 1; i < n; ++i) {
        const double d = std::fabs(i * h - p);
        if (d < bd) {
            bd = d;
            best = i;
        }
    }
    return best;
}

// True for memory the host cannot read directly (device memory); host memory, pageable or pinned,
// is read in place.
bool device_memory(const void* p)
{
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;  // not known to HIP: pageable host memory
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeArray;
}

// A small input array (coordinates, spectra, time axes) on the host. Device memory is copied with
// a synchronous copy, which waits for the work queued before it; host memory is read without
// waiting.
template <class T>
std::vector<T> to_host(const void* src, size_t n)
{
    std::vector<T> h(n);
    if (n == 0) return h;
    if (device_memory(src)) HIPCHK(hipMemcpy(h.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
    else std::memcpy(h.data(), src, n * sizeof(T));
    return h;
}

std::vector<int2> node_ids(const fd2d_carray* pts, const Geometry& g, const char* what)
{
    std::vector<int2> ids;
    const int64_t n = numel(pts) ? pts->dims[0] : 0;
    if (n == 0) return ids;
    if (pts->dims[1] != 2 || pts->dims[2] != 1) raise(FD2D_EINVAL, "%s must be an n x 2 array", what);
    const std::vector<double> h = to_host<double>(pts->data, (size_t)n * 2);
    ids.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        ids[(size_t)i].x = nearest_node(h[(size_t)i], g.dx, g.nx);
        ids[(size_t)i].y = nearest_node(h[(size_t)(i + n)], g.dz, g.nz);
    }
    return ids;
}

// interpol_matrix of code/run_forward2_correlation.m:121-165 with the band limits of
// code/propagation/integration_limits.m:2-14 (K x max(nb,1), column-major).
std::vector<double> interpol_matrix(int K, int nb)
{
    if (nb == 0) return std::vector<double>((size_t)K, 1.0);
    std::vector<double> I((size_t)K * nb, 0.0);
    const double step = (double)K / (double)nb;
    for (int k = 1; k <= K; ++k) {
        int ib = -1, lo = 0, hi = 0;
        for (int b = 1; b <= nb; ++b) {
            const int l = (int)std::floor((b - 1) * step + 1), h = (int)std::floor(b * step);
            if (k >= l && k <= h) {
                ib = b;
                lo = l;
                hi = h;
                break;
            }
        }
        if (ib < 0) raise(FD2D_EINVAL, "frequency %d lies in no band", k);
        const int ni = hi - lo + 1;
        const int below = ib == 1 ? 0 : ib - 1;
        const int above = ib == 1 ? 2 : (ib == nb ? 0 : ib + 1);
        const double i = (double)(k - lo + 1), half = ni / 2.0;
        auto at = [&](int kk, int bb) -> double& { return I[(size_t)(kk - 1) + (size_t)K * (bb - 1)]; };
        if (i < half) {
            at(k, ib) = 1 - (half - i) / ni;
            if (below) at(k, below) = (half - i) / ni;
        } else if (i == half) {
            at(k, ib) = 1;
        } else {
            at(k, ib) = 1 - (i - half) / ni;
            if (above) at(k, above) = (i - half) / ni;
        }
    }
    return I;
}

// Separable Gaussian taper of code/propagation/init_absbound.m:16-30: absbound(i,j) = ax(i)*az(j)
// (FP64, as the reference builds it).
void taper_d(const fd2d_config* cfg, const Geometry& g, std::vector<double>& ax, std::vector<double>& az)
{
    const double w = cfg->absorb_width, Lx = cfg->Lx, Lz = cfg->Lz;
    auto lo_side = [w](double X) { return (double)(X > w) + std::exp(-(X - w) * (X - w) / ((2 * w) * (2 * w))) * (double)(X <= w); };
    auto hi_side = [w](double X, double L) {
        return (double)(X < (L - w)) + std::exp(-(X - (L - w)) * (X - (L - w)) / ((2 * w) * (2 * w))) * (double)(X >= (L - w));
    };
    ax.resize(g.nx);
    az.resize(g.nz);
    for (int i = 0; i < g.nx; ++i) {
        const double X = i * g.dx;
        double a = 1.0;
        if (cfg->absorb_left == 1) a *= lo_side(X);
        if (cfg->absorb_right == 1) a *= hi_side(X, Lx);
        ax[i] = a;
    }
    for (int j = 0; j < g.nz; ++j) {
        const double Z = j * g.dz;
        double a = 1.0;
        if (cfg->absorb_bottom == 1) a *= lo_side(Z);
        if (cfg->absorb_top == 1) a *= hi_side(Z, Lz);
        az[j] = a;
    }
}

void taper(const fd2d_config* cfg, const Geometry& g, std::vector<float>& ax, std::vector<float>& az)
{
    std::vector<double> axd, azd;
    taper_d(cfg, g, axd, azd);
    ax.assign(axd.begin(), axd.end());
    az.assign(azd.begin(), azd.end());
}

// Copies a host table to a new device buffer on stream st; the buffer keeps the host copy the
// transfer reads until it is released, so the caller's vector may go away at once.
template <class T>
void upload(DBuf<T>& d, const std::vector<T>& h, hipStream_t st)
{
    d.alloc(h.size());
    if (h.empty()) return;
    const unsigned char* b = reinterpret_cast<const unsigned char*>(h.data());
    d.stage.assign(b, b + h.size() * sizeof(T));
    HIPCHK(hipMemcpyAsync(d.p, d.stage.data(), d.stage.size(), hipMemcpyHostToDevice, st));
}

// Phase table e^{sign i w_k t_j} * scale for the listed times, rows j, columns k.
std::vector<double2> phase_table(const fd2d_config* cfg, const std::vector<double>& t, double sign, double scale)
{
    const int K = cfg->n_sample;
    std::vector<double2> c(t.size() * (size_t)K);
    for (size_t j = 0; j < t.size(); ++j)
        for (int k = 0; k < K; ++k) {
            const double ph = sign * cfg->w_sample[k] * t[j];
            c[j * K + k] = make_double2(scale * std::cos(ph), scale * std::sin(ph));
        }
    return c;
}

const double kInvSqrt2Pi = 1.0 / std::sqrt(2.0 * M_PI);

// Time-stepping state of one solve on one stream.
struct Solver {
    const fd2d_config* cfg;
    Geometry g;
    hipStream_t st;
    bool with_u;
    // displacement form (correlation / adjoint runs at FD order 6 or 8 without Born terms): the pool
    // holds v (FP32) and u (FP64) only, sigma is evaluated from u inside the step
    bool uform;
    PaddedLayout lay;
    DBuf<float> fields;  // two sets of v, sxy, szy (, u); displacement form: v0, v1, then u0, u1 in FP64
    static constexpr size_t kBinfo = 64;  // words after the two coefficient arrays: binfo
    DBuf<float> coefs;   // bv, dtmu, binfo
    DBuf<float> ax_d, az_d, axp_d, azp_d;
    float *v[2], *sxy[2], *szy[2], *u[2];
    double* u64[2] = {nullptr, nullptr};  // displacement form: origin pointers of the FP64 u arrays
    int o_u64[2] = {0, 0};                // ... and their FP32-element offsets in the pool
    int cur = 0;
    int tiles_x = 0, tiles_z = 0, n_tiles = 0;
    DBuf<int> ps_off_d, rec_off_d;
    DBuf<int2> ps_ent_d, rec_ent_d;
    StepArgs a{};
    PairArgs pr{};  // the second step of a two-step launch

    Solver(const fd2d_config* c, const Geometry& geo, hipStream_t s, bool with_u_, bool uform_ = false)
        : cfg(c), g(geo), st(s), with_u(with_u_), uform(uform_)
    {
        lay = PaddedLayout::make(g.nx, g.nz);
        const size_t E = lay.elems(), o = lay.origin();
        // buffer offsets are 32-bit byte offsets into the field pool (two sets of nf arrays)
        if (E * sizeof(float) * 8 >= ((size_t)1 << 32))
            raise(FD2D_EUNSUPPORTED, "nx x nz = %d x %d: the padded fields exceed 4 GiB", g.nx, g.nz);
        // FP32 arrays per set: v, sxy, szy (, u); displacement form: v and the two halves of FP64 u
        const int nf = uform ? 3 : with_u ? 4 : 3;
        fields.alloc(E * nf * 2);
        HIPCHK(hipMemsetAsync(fields.p, 0, E * nf * 2 * sizeof(float), st));
        coefs.alloc(E * 2 + kBinfo);
        HIPCHK(hipMemsetAsync(coefs.p, 0, (E * 2 + kBinfo) * sizeof(float), st));
        std::vector<float> ax, az;
        taper(cfg, g, ax, az);
        upload(ax_d, ax, st);
        upload(az_d, az, st);
        std::vector<float> axp((size_t)lay.pitch, 0.f), azp((size_t)lay.rows, 0.f);
        for (int i = 0; i < g.nx; ++i) axp[(size_t)i + kGhostX] = ax[(size_t)i];
        for (int j = 0; j < g.nz; ++j) azp[(size_t)j + kGhost] = az[(size_t)j];
        upload(axp_d, axp, st);
        upload(azp_d, azp, st);
        for (int b = 0; b < 2; ++b) {
            if (uform) {
                v[b] = fields.p + (size_t)b * E + o;
                sxy[b] = szy[b] = u[b] = nullptr;
                o_u64[b] = (int)(2 * E + (size_t)b * 2 * E);
                u64[b] = (double*)(fields.p + o_u64[b]) + o;
                continue;
            }
            float* base = fields.p + (size_t)b * nf * E;
            v[b] = base + o;
            sxy[b] = base + E + o;
            szy[b] = base + 2 * E + o;
            u[b] = with_u ? base + 3 * E + o : nullptr;
        }
        a.bv = coefs.p + o;
        a.dtmu = coefs.p + E + o;
        a.binfo = coefs.p + 2 * E;  // past the pool's buffer range (cbytes)
        a.ax = ax_d.p;
        a.az = az_d.p;
        a.axp = axp_d.p;
        a.azp = azp_d.p;
        a.fbase = fields.p;
        a.fbytes = (unsigned)(E * nf * 2 * sizeof(float));
        a.cbase = coefs.p;
        a.cbytes = (unsigned)(E * 2 * sizeof(float));
        a.o_bv = 0;
        a.o_dtmu = (int)E;
        a.compact_bytes = (unsigned)(g.P * sizeof(float));
        a.nx = g.nx;
        a.nz = g.nz;
        a.pitch = lay.pitch;
        tiles_x = (g.nx + kTX - 1) / kTX;
        tiles_z = (g.nz + kTileZ - 1) / kTileZ;
        n_tiles = tiles_x * tiles_z;
        a.tiles_x = tiles_x;
        a.n_tiles = n_tiles;
        a.tiles_per_xcd = (n_tiles + kXCD - 1) / kXCD;
        a.origin = (int)o;
        a.padded_bytes = (unsigned)(E * sizeof(float));
        for (int m = 0; m < 4; ++m) {
            a.cx[m] = (float)(kCoef[g.R][m] / g.dx);
            a.cz[m] = (float)(kCoef[g.R][m] / g.dz);
        }
        a.dt = (float)g.dt;
        a.dt64 = g.dt;
    }
    // dt/rho and dt*mu (velocity-stress form) or mu (displacement form)
    void material(const double* mu, const double* rho)
    {
        launch_material(mu, rho, coefs.p + lay.origin(), coefs.p + lay.elems() + lay.origin(), g.nx, g.nz,
                        lay.pitch, g.dt, uform ? 1.0 : g.dt, coefs.p + 2 * lay.elems(), st);
    }
    // the FP32 part of the pool a finiteness check reads (the FP64 halves of u are not FP32 values)
    const float* check_ptr() const { return fields.p; }
    size_t check_count() const { return uform ? 2 * lay.elems() : fields.n; }
    // Point sources as per-tile lists over every node whose velocity the tile computes in a one- or
    // two-step launch (the tile plus its (3R-2, 3R-1) halo), in ascending source order so every
    // copy of a halo node sums identically.
    void set_sources(const std::vector<int2>& nodes)
    {
        if (nodes.empty()) return;
        const int lo = pair_halo_lo(g.R), hi = pair_halo_hi(g.R);
        std::vector<std::vector<int2>> per(n_tiles);
        for (size_t q = 0; q < nodes.size(); ++q) {
            const int x = nodes[q].x, z = nodes[q].y;
            for (int t = 0; t < n_tiles; ++t) {
                const int i0 = (t % tiles_x) * kTX, j0 = (t / tiles_x) * kTileZ;
                if (x < i0 - lo || x > i0 + kTX - 1 + hi) continue;
                if (z < j0 - lo || z > j0 + kTileZ - 1 + hi) continue;
                per[t].push_back(make_int2(node_key_host(x, z), (int)q));
            }
        }
        csr(per, ps_off_d, ps_ent_d);
    }
    // Receivers as per-tile lists over its own nodes.
    void set_receivers(const std::vector<int2>& nodes)
    {
        if (nodes.empty()) return;
        std::vector<std::vector<int2>> per(n_tiles);
        for (size_t q = 0; q < nodes.size(); ++q) {
            const int x = nodes[q].x, z = nodes[q].y;
            per[(z / kTileZ) * tiles_x + x / kTX].push_back(make_int2(node_key_host(x, z), (int)q));
        }
        csr(per, rec_off_d, rec_ent_d);
    }
    void csr(const std::vector<std::vector<int2>>& per, DBuf<int>& off_d, DBuf<int2>& ent_d)
    {
        std::vector<int> off(per.size() + 1, 0);
        std::vector<int2> ent;
        for (size_t t = 0; t < per.size(); ++t) {
            ent.insert(ent.end(), per[t].begin(), per[t].end());
            off[t + 1] = (int)ent.size();
        }
        upload(off_d, off, st);
        upload(ent_d, ent, st);
    }
    void clear_step()
    {
        a.v_in = v[cur];
        a.sxy_in = sxy[cur];
        a.szy_in = szy[cur];
        a.u_in = u[cur];
        a.v_out = v[1 - cur];
        a.sxy_out = sxy[1 - cur];
        a.szy_out = szy[1 - cur];
        a.u_out = u[1 - cur];
        auto off = [&](const float* p) { return p ? (int)(p - (const float*)fields.p - (ptrdiff_t)lay.origin()) : 0; };
        a.o_v_in = off(a.v_in);
        a.o_sxy_in = off(a.sxy_in);
        a.o_szy_in = off(a.szy_in);
        a.o_u_in = off(a.u_in);
        a.o_v_out = off(a.v_out);
        a.o_sxy_out = off(a.sxy_out);
        a.o_szy_out = off(a.szy_out);
        a.o_u_out = off(a.u_out);
        if (uform) {
            a.o_u_in = o_u64[cur];
            a.o_u_out = o_u64[1 - cur];
        }
        a.ps_off = nullptr;
        a.ps_ent = nullptr;
        a.ps_amp = nullptr;
        a.ps_const = 0.f;
        a.syn = nullptr;
        a.snap = nullptr;
        a.dsamp = nullptr;
        a.rec_off = nullptr;
        a.rec_ent = nullptr;
        a.seis_col = nullptr;
        a.born_dmu = a.born_in0 = a.born_in1 = nullptr;
        a.kfwd = nullptr;
        pr.ps_amp2 = nullptr;
        pr.ps_const2 = 0.f;
        pr.syn2 = nullptr;
        pr.snap2 = nullptr;
        pr.dsamp2 = nullptr;
        pr.seis_col2 = nullptr;
    }
    // The per-step fields of step `sub` of a launch (0: a single step or the first of a pair).
    struct Slots {
        const float** ps_amp;
        float* ps_const;
        const float** syn;
        float** snap;
        float** dsamp;
        double** seis;
    };
    Slots slots(int sub)
    {
        if (sub == 0) return {&a.ps_amp, &a.ps_const, &a.syn, &a.snap, &a.dsamp, &a.seis_col};
        return {&pr.ps_amp2, &pr.ps_const2, &pr.syn2, &pr.snap2, &pr.dsamp2, &pr.seis_col2};
    }
    void point_sources(int sub, const float* amp, float cst)
    {
        a.ps_off = ps_off_d.p;
        a.ps_ent = ps_ent_d.p;
        const Slots sl = slots(sub);
        *sl.ps_amp = amp;
        *sl.ps_const = cst;
    }
    void receivers(int sub, double* col)
    {
        a.rec_off = rec_off_d.p;
        a.rec_ent = rec_ent_d.p;
        *slots(sub).seis = col;
    }
    void born_green(const float* dtdmu, const float* gin)
    {
        launch_born_green(g.R, a, sxy[cur], szy[cur], dtdmu, gin, st);
    }
    void step()
    {
        if (uform) launch_ustep(g.R, a, st);
        else launch_step(g.R, with_u, a, st);
        cur ^= 1;
    }
    // steps n and n+1 in one launch (the state buffers swap once, as after one step)
    void step_pair()
    {
        pr.a = a;
        launch_pair(g.R, with_u, pr, st);
        cur ^= 1;
    }
    // structure-kernel product with the displacement the last step produced
    void kmu(const float* fwd, double* kmu_acc, double scale)
    {
        if (uform) launch_ukmu(g.R, a, u64[cur], fwd, kmu_acc, scale, st);
        else launch_kmu(g.R, a, u[cur], fwd, kmu_acc, scale, st);
    }
};

// Blocked frequency-domain helpers: time samples are staged in FP32 slabs of up to kSpecSlab slices
// and contracted on the matrix cores (csrc/spectral.hip). The slab length depends only on the grid
// and the sample count (so a solve sums in the same order wherever and whenever it runs): kSpecSlab
// samples, halved while a slab would exceed kSlabBytes (grids above about 2900^2).
int spec_cols(int K) { return (2 * K + 255) / 256 * 256; }

constexpr size_t kSlabBytes = size_t(32) << 30;

int slab_len(size_t P, size_t n_samples)
{
    size_t T = std::min<size_t>(kSpecSlab, std::max<size_t>(n_samples, 1));
    while (T > 64 && T * P * sizeof(float) > kSlabBytes) T /= 2;
    // several slabs: each starts on a 16-sample boundary (the phase tables are read 8 samples at a time)
    if (T < n_samples) T = std::max<size_t>(16, T / 16 * 16);
    return (int)T;
}

// The three bfloat16 parts of a phase table, each [rows][cols] (uint16), zero where not set.
struct SplitTable {
    size_t rows = 0, cols = 0;
    std::vector<unsigned short> h, m, l;
    void resize(size_t r, size_t c)
    {
        rows = r;
        cols = c;
        h.assign(r * c, 0);
        m.assign(r * c, 0);
        l.assign(r * c, 0);
    }
    void set(size_t r, size_t c, double v) { split_bf16((float)v, h[r * cols + c], m[r * cols + c], l[r * cols + c]); }
};

// Split phase tables on the device, cached per device: the solves of one configuration (the
// stations of an inversion, the requests of a served model) share a table instead of recomputing
// its sines and cosines and uploading it again, which costs milliseconds of host time per solve.
// An entry is keyed by everything the table is computed from; it is uploaded on the stream of the
// solve that first needs it, and later users wait for that upload through its event. Entries no
// solve holds are freed least recently used first once the cache exceeds kTableCacheBytes.
struct DeviceTable {
    unsigned short* part[3] = {nullptr, nullptr, nullptr};
    std::vector<unsigned short> host[3];  // staging of the asynchronous upload
    size_t n = 0;                         // elements per part
    hipEvent_t ready = nullptr;     // the upload is complete; null once it is known to be
    hipStream_t ready_st = nullptr;  // the stream of the upload
    ~DeviceTable()
    {
        for (unsigned short* q : part)
            if (q) (void)hipFree(q);
        if (ready) (void)hipEventDestroy(ready);
    }
};

constexpr size_t kTableCacheBytes = size_t(256) << 20;
std::mutex g_table_mu;
struct TableCache {
    std::vector<std::pair<std::string, std::shared_ptr<DeviceTable>>> lru;  // most recent last
    size_t bytes = 0;
};
// never destroyed: at process exit the HIP runtime may already be gone, and the memory goes with it
TableCache* const g_tables = new TableCache[64];

template <class T>
void key_add(std::string& k, const T* v, size_t n)
{
    k.append(reinterpret_cast<const char*>(v), n * sizeof(T));
}

// The table for `key`, built by `build` (which fills a SplitTable) when it is not cached.
template <class F>
std::shared_ptr<const DeviceTable> cached_table(const std::string& key, hipStream_t st, F&& build)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    TableCache& c = g_tables[dev & 63];
    std::lock_guard<std::mutex> lk(g_table_mu);
    for (size_t q = 0; q < c.lru.size(); ++q)
        if (c.lru[q].first == key) {
            std::shared_ptr<DeviceTable> t = c.lru[q].second;
            c.lru.erase(c.lru.begin() + (long)q);
            c.lru.emplace_back(key, t);
            if (t->ready) HIPCHK(hipStreamWaitEvent(st, t->ready, 0));
            return t;
        }
    SplitTable e;
    build(e);
    auto t = std::make_shared<DeviceTable>();
    t->n = e.h.size();
    t->host[0] = std::move(e.h);
    t->host[1] = std::move(e.m);
    t->host[2] = std::move(e.l);
    for (int q = 0; q < 3; ++q) {
        if (hipMalloc(&t->part[q], std::max<size_t>(t->n, 1) * sizeof(unsigned short)) != hipSuccess) {
            (void)hipGetLastError();
            t->part[q] = nullptr;
            raise(FD2D_ENOMEM, "device allocation of a phase table failed");
        }
        if (t->n)
            HIPCHK(hipMemcpyAsync(t->part[q], t->host[q].data(), t->n * sizeof(unsigned short), hipMemcpyHostToDevice,
                                  st));
    }
    HIPCHK(hipEventCreateWithFlags(&t->ready, hipEventDisableTiming));
    HIPCHK(hipEventRecord(t->ready, st));
    t->ready_st = st;
    c.lru.emplace_back(key, t);
    c.bytes += 3 * t->n * sizeof(unsigned short);
    for (size_t q = 0; c.bytes > kTableCacheBytes && q + 1 < c.lru.size();) {
        if (c.lru[q].second.use_count() == 1) {  // no solve holds it, so no queued work reads it
            c.bytes -= 3 * c.lru[q].second->n * sizeof(unsigned short);
            c.lru.erase(c.lru.begin() + (long)q);
        } else {
            ++q;
        }
    }
    return t;
}

// Before the stream s is destroyed: the cached device blocks and tables whose events were
// recorded on s lose those events (the work they mark is waited for first).
void detach_stream(hipStream_t s)
{
    for (BlockCache& c : g_cache) c.detach(s);
    std::lock_guard<std::mutex> lk(g_table_mu);
    for (int d = 0; d < 64; ++d)
        for (auto& kv : g_tables[d].lru) {
            DeviceTable& t = *kv.second;
            if (t.ready && t.ready_st == s) {
                if (hipEventSynchronize(t.ready) != hipSuccess) (void)hipGetLastError();
                (void)hipEventDestroy(t.ready);
                t.ready = nullptr;
                t.ready_st = nullptr;
            }
        }
}

// Key of the phase table e^{sign i w_k t_j} * scale over the times ts, in layout `kind` with slab
// length T and C_pad columns.
std::string table_key(char kind, const fd2d_config* cfg, const std::vector<double>& ts, double sign, double scale,
                      int T, int C_pad)
{
    std::string k(1, kind);
    const int32_t ints[4] = {T, C_pad, cfg->n_sample, (int32_t)ts.size()};
    const double dbl[2] = {sign, scale};
    key_add(k, ints, 4);
    key_add(k, dbl, 2);
    key_add(k, ts.data(), ts.size());
    key_add(k, cfg->w_sample, (size_t)cfg->n_sample);
    return k;
}

struct DftAccumulator {
    double2* acc = nullptr;
    DBuf<double2> comp;  // Neumaier compensation of acc: the sum is acc + comp (series of several slabs)
    size_t P = 0;
    int K = 0, C_pad = 0, Jt = 0;
    DBuf<float> slab;
    // [C_pad][Jt]: parts of (Re, Im) of each sample's phase factor
    std::shared_ptr<const DeviceTable> tab;
    int cnt = 0, done = 0, T = 0, n_total = 0;
    // where the last slab writes the finished sums: out = scale * (conj ? conj(sum) : sum)
    double2* out = nullptr;
    double out_scale = 1.0;
    int out_conj = 0;
    hipStream_t st = nullptr;
    // phase factors e^{sign i w_k t_j} * scale at the sample times ts; the finished series goes to
    // out_ (which may be acc_) as out_scale_ * (conj ? conj(sum) : sum)
    void init(double2* acc_, size_t P_, const fd2d_config* cfg, const std::vector<double>& ts, double sign,
              double scale, double2* out_, double out_scale_, int out_conj_, hipStream_t s)
    {
        acc = acc_;
        P = P_;
        K = cfg->n_sample;
        st = s;
        out = out_;
        out_scale = out_scale_;
        out_conj = out_conj_;
        C_pad = spec_cols(K);
        const size_t n_samples = ts.size();
        n_total = (int)n_samples;
        T = slab_len(P, n_samples);  // one slab, or a multiple of 16 samples
        Jt = (int)((n_samples + 15) / 16 * 16);
        slab.alloc((size_t)T * P);
        tab = cached_table(table_key('D', cfg, ts, sign, scale, T, C_pad), st, [&](SplitTable& e) {
            const std::vector<double2> table = phase_table(cfg, ts, sign, scale);
            e.resize(C_pad, Jt);
            for (size_t j = 0; j < n_samples; ++j)
                for (int k = 0; k < K; ++k) {
                    e.set(2 * k, j, table[j * K + k].x);
                    e.set(2 * k + 1, j, table[j * K + k].y);
                }
        });
        // the first slab starts the sums without reading them, and a series of one slab keeps no
        // compensation terms between slabs
        if (n_total > T) comp.alloc(P * (size_t)K);
    }
    float* next_slot() { return slab.p + (size_t)cnt * P; }
    void commit()
    {
        if (++cnt == T) flush();
    }
    // contract the staged samples; the slab that completes the series writes the finished sums
    void flush()
    {
        if (cnt == 0) return;
        const bool last = done + cnt >= n_total;
        launch_dft_bf16(acc, comp.p, slab.p, tab->part[0] + done, tab->part[1] + done, tab->part[2] + done, P, K, C_pad,
                        Jt, cnt, done == 0 ? 1 : 0, last ? out : nullptr, out_scale, out_conj, st);
        done += cnt;
        cnt = 0;
    }
    // after the series' last sample: the remaining samples, or zeros for a series without samples
    void finish()
    {
        flush();
        if (n_total == 0 && K > 0) HIPCHK(hipMemsetAsync(out, 0, P * (size_t)K * sizeof(double2), st));
    }
};

struct Synthesiser {
    size_t P = 0;
    int K = 0, C_pad = 0, T = 0, J_pad = 0, n_total = 0;
    DBuf<float> slab;
    DBuf<float> a32;  // [C_pad][P]: (Re, Im) rows of the source spectrum, FP32
    // per slab [J_pad][C_pad]: parts of (Re, -Im) of the phase factors
    std::shared_ptr<const DeviceTable> tab;
    hipStream_t st = nullptr;
    // the FP32 spectrum rows, for the caller to fill (launch_source_spectrum) before init
    float* spectrum_rows(size_t P_, int K_)
    {
        P = P_;
        K = K_;
        C_pad = spec_cols(K);
        a32.alloc((size_t)C_pad * P);
        return a32.p;
    }
    // phase factors e^{i w_k t_j} * scale at the slice times ts
    void init(const fd2d_config* cfg, const std::vector<double>& ts, double scale, hipStream_t s)
    {
        st = s;
        const int n_samples = (int)ts.size();
        n_total = n_samples;
        T = slab_len(P, (size_t)std::max(n_samples, 1));
        J_pad = (T + 63) / 64 * 64;
        slab.alloc((size_t)T * P);
        tab = cached_table(table_key('S', cfg, ts, 1.0, scale, T, C_pad), st, [&](SplitTable& e) {
            const std::vector<double2> table = phase_table(cfg, ts, 1.0, scale);
            const int n_slabs = (std::max(n_samples, 1) + T - 1) / T;
            e.resize((size_t)n_slabs * J_pad, C_pad);
            for (int idx = 0; idx < n_samples; ++idx) {
                const size_t row = (size_t)(idx / T) * J_pad + (size_t)(idx % T);
                for (int k = 0; k < K; ++k) {
                    e.set(row, 2 * k, table[(size_t)idx * K + k].x);
                    e.set(row, 2 * k + 1, -table[(size_t)idx * K + k].y);
                }
            }
        });
    }
    // Source slice of the idx-th sampled step (0-based); synthesises a new block when needed.
    const float* slice(int idx)
    {
        const int slot = idx % T;
        if (slot == 0) {
            const size_t o = (size_t)(idx / T) * J_pad * C_pad;
            launch_synth_bf16(slab.p, a32.p, tab->part[0] + o, tab->part[1] + o, tab->part[2] + o, P, C_pad,
                              std::min(T, n_total - idx), st);
        }
        return slab.p + (size_t)slot * P;
    }
};

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------------------------
// Step-kernel timing (fd2d_timing_enable / fd2d_timing_read)
// ---------------------------------------------------------------------------------------------
std::mutex g_timing_mu;
int g_timing_every = 0;
std::atomic<int> g_steps_per_launch{0};  // fd2d_set_steps_per_launch: 0 = tuned per shape
fd2d_timing g_timing{};

struct StepTimer {
    int solver, every;
    size_t P;
    hipStream_t st;
    std::vector<hipEvent_t> ev;  // pairs: before and after a step launch
    std::vector<int> steps;      // time steps each timed launch advanced
    StepTimer(int solver_, size_t P_, hipStream_t s) : solver(solver_), P(P_), st(s)
    {
        std::lock_guard<std::mutex> lk(g_timing_mu);
        every = g_timing_every;
    }
    // a launch of `k` steps from step n is timed when one of its steps is a multiple of `every`
    bool on(int n, int k) const { return every > 0 && (n + k - 1) / every > (n - 1) / every; }
    // Timing-only events: without the system-scope fence of a default event, whose cache
    // write-back and invalidation would be counted in the timed launch and slow the next one.
    // collect() runs after the stream has been synchronised, so the times need no fence.
    void record()
    {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        HIPCHK(hipEventRecord(e, st));
        ev.push_back(e);
    }
    void begin(int n, int k)
    {
        if (!on(n, k)) return;
        record();
        steps.push_back(k);
    }
    void end(int n, int k)
    {
        if (on(n, k)) record();
    }
    // Call after the stream has been synchronised.
    void collect()
    {
        double ms = 0;
        int64_t cnt = 0, nsteps = 0;
        for (size_t q = 0; q + 1 < ev.size(); q += 2) {
            float e = 0;
            HIPCHK(hipEventElapsedTime(&e, ev[q], ev[q + 1]));
            ms += e;
            ++cnt;
            nsteps += steps[q / 2];
        }
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
        steps.clear();
        if (cnt) {
            std::lock_guard<std::mutex> lk(g_timing_mu);
            g_timing.launches[solver - 1][0] += cnt;
            g_timing.launches[solver - 1][1] += nsteps;
            g_timing.ms[solver - 1][0] += ms;
            g_timing.points[solver - 1] = (int64_t)P;
        }
    }
    ~StepTimer()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

void check_out(const fd2d_marray* o, int64_t d0, int64_t d1, int64_t d2, const char* what)
{
    if (!o || !o->data) return;
    if (o->dims[0] * o->dims[1] * o->dims[2] != d0 * d1 * d2)
        raise(FD2D_EINVAL, "output %s must hold %lld x %lld x %lld elements", what, (long long)d0, (long long)d1,
              (long long)d2);
}

int64_t green_n_fw(const Geometry& g)
{
    int64_t n = 0;
    for (int64_t q = g.nt; q <= 2 * (int64_t)g.nt - 1; ++q)
        if (q % g.f == 0) ++n;
    return n;
}

// Correlation and adjoint runs use the displacement form at FD orders 6 and 8 (its 32 B per step
// against 40 B; at orders 2 and 4 two-step launches of the velocity-stress form move fewer bytes
// still) and when there are no Born terms (whose stress terms have no displacement counterpart).
bool displacement_form(const Geometry& g, bool born) { return g.R >= 3 && !born; }

// =============================================================================================
// Solves as step sequences. A Job holds one solve's set-up (the prologue of the *_device entry
// points), fills the per-step fields of step n, does the work after a launch and finishes the
// solve. run_job() advances one job alone (one or two steps per launch); the batch engine below
// advances many jobs with grouped launches.
// =============================================================================================
struct Job {
    const fd2d_config* cfg = nullptr;
    Geometry g{};
    hipStream_t st = nullptr;
    std::unique_ptr<Solver> S;
    int solver = 0, nsteps = 0, next = 1;
    bool born = false;      // Born terms: one step per launch
    bool records = false;   // writes seismograms (receiver lists are used)
    virtual ~Job() = default;
    virtual void plan(int n, int sub) = 0;  // fields of step n as step `sub` of the launch
    virtual bool pairable(int n) const = 0;  // may steps n, n+1 share a two-step launch
    virtual void after(int n, int k) = 0;   // after the launch that advanced steps n .. n+k-1
    virtual void finish() = 0;              // after the last step
    // For a step launched on its own: move the structure-kernel product that plan() scheduled
    // after the launch into the step kernel itself.
    virtual void fuse_products() {}
    // whether step n (not yet planned) carries a structure-kernel product
    virtual bool product_at(int) const { return false; }
    // StepArgs with the fields that do not change from step to step, for ping-pong direction `parity`
    StepArgs static_args(int parity)
    {
        const int keep = S->cur;
        S->cur = parity;
        S->clear_step();
        StepArgs s = S->a;
        S->cur = keep;
        S->clear_step();
        if (S->ps_ent_d.n) {
            s.ps_off = S->ps_off_d.p;
            s.ps_ent = S->ps_ent_d.p;
        }
        if (records && S->rec_ent_d.n) {
            s.rec_off = S->rec_off_d.p;
            s.rec_ent = S->rec_ent_d.p;
        }
        return s;
    }
};

// Launch-shape autotuner. With fd2d_set_steps_per_launch(0) (the default) the first solve of each
// shape (grid, FD order, field set) times its own early launches both ways -- kTuneLaunches single
// steps, then kTuneLaunches two-step launches -- and the faster shape is kept for the rest of that
// solve and every later solve of the shape. Both shapes give the same results bit for bit, so the
// choice changes only the time. Solves with Born terms, or too short to time, use single steps.
constexpr int kTuneLaunches = 24;
constexpr int kTuneWarm = 12;
std::mutex g_tune_mu;
std::map<std::array<int, 4>, int> g_tuned;  // (nx, nz, R, with_u) -> steps per launch

std::array<int, 4> shape_key(const Job& j) { return {j.g.nx, j.g.nz, j.g.R, j.S->with_u ? 1 : 0}; }

// field set of a solve: 0 Green (v, sigma), 1 with displacement (v, sigma, u), 2 displacement form (v, FP64 u)
int field_kind(const Job& j) { return j.S->uform ? 2 : j.S->with_u ? 1 : 0; }

int tuned_steps(const Job& j)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(shape_key(j));
    return it == g_tuned.end() ? 0 : it->second;
}

void record_tuned(const Job& j, int steps)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[shape_key(j)] = steps;
}

// steps per launch of solve j: 1 or 2, or 0 when this solve should time both
int launch_steps(const Job& j)
{
    const int mode = g_steps_per_launch.load();
    if (j.born || j.S->uform) return 1;
    if (mode == 1 || mode == 2) return mode;
    const int t = tuned_steps(j);
    if (t) return t;
    return j.nsteps >= 8 * kTuneLaunches + kTuneWarm ? 0 : 1;
}

void run_job(Job& j)
{
    StepTimer timer(j.solver, j.g.P, j.st);
    int steps = launch_steps(j);
    // tuning: phase 0 = single steps, phase 1 = pairs, each over kTuneLaunches launches, after
    // kTuneWarm launches that hold the solve's one-off work (the first synthesised source block)
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int phase = steps == 0 ? -2 : -1, launches = 0, advanced[2] = {0, 0};
    auto mark = [&](int q) {
        HIPCHK(hipEventCreate(&ev[q]));
        HIPCHK(hipEventRecord(ev[q], j.st));
    };
    for (int n = 1; n <= j.nsteps;) {
        if (phase == -2 && n > kTuneWarm) {
            mark(0);
            phase = 0;
        }
        const bool pairs = phase == 1 || (phase == -1 && steps == 2);
        const int k = (pairs && n + 1 <= j.nsteps && j.pairable(n)) ? 2 : 1;
        j.S->clear_step();
        for (int q = 0; q < k; ++q) j.plan(n + q, q);
        if (k == 1) j.fuse_products();
        timer.begin(n, k);
        if (k == 2) j.S->step_pair();
        else j.S->step();
        timer.end(n, k);
        j.after(n, k);
        n += k;
        if (phase >= 0) {
            advanced[phase] += k;
            if (++launches == kTuneLaunches) {
                mark(2 * phase + 1);
                launches = 0;
                if (phase == 0) {
                    mark(2);
                    phase = 1;
                } else {
                    HIPCHK(hipEventSynchronize(ev[3]));
                    float t0 = 0.f, t1 = 0.f;
                    HIPCHK(hipEventElapsedTime(&t0, ev[0], ev[1]));
                    HIPCHK(hipEventElapsedTime(&t1, ev[2], ev[3]));
                    // pairs only when clearly faster per time step
                    steps = (t1 / advanced[1] < 0.97f * (t0 / advanced[0])) ? 2 : 1;
                    record_tuned(j, steps);
                    phase = -1;
                }
            }
        }
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    j.finish();
    // the outputs are written either way; a NaN or Inf left in the fields is reported
    DBuf<int> chk(kCheckBlocks + 1);
    launch_nonfinite(j.S->check_ptr(), j.S->check_count(), chk.p, chk.p + kCheckBlocks, j.st);
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, chk.p + kCheckBlocks, sizeof bad, hipMemcpyDeviceToHost, j.st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(j.st));
    timer.collect();
    if (bad)
        raise(FD2D_ENONFINITE, "the wavefield became non-finite (NaN or Inf) within %d steps: dt = %g may exceed the "
              "stability (CFL) limit of the order-%d stencil at dx = %g, dz = %g", j.nsteps, j.g.dt, 2 * j.g.R,
              j.g.dx, j.g.dz);
}

// =============================================================================================
// run_forward1_green (code/run_forward1_green.m)
// =============================================================================================
struct GreenJob : Job {
    double mode;
    const fd2d_carray* G_in;
    std::vector<int2> sid, rid;
    int64_t nr = 0, i_fw_out = 0;
    double2* Gf = nullptr;
    float* Gout = nullptr;
    double* sg = nullptr;
    DBuf<float> dtdmu, ndtdmu;  // dt * dmu and -dt * dmu (FP32)
    DftAccumulator dft;

    GreenJob(const fd2d_config* c, const fd2d_carray* mu, const fd2d_carray* rho, const fd2d_carray* src,
             const fd2d_carray* rec, double mode_, const fd2d_carray* dmu, const fd2d_carray* G_in_,
             fd2d_marray* G_fft, fd2d_marray* G_out, fd2d_marray* seis, hipStream_t s)
        : mode(mode_), G_in(G_in_)
    {
        cfg = c;
        st = s;
        solver = FD2D_GREEN;
        g = geometry(cfg);
        const int64_t P = (int64_t)g.P;
        if (numel(mu) != P || numel(rho) != P)
            raise(FD2D_EINVAL, "mu and rho must have nx*nz = %lld elements", (long long)P);
        born = numel(dmu) != 0;
        if (born && numel(dmu) != P) raise(FD2D_EINVAL, "dmu must have nx*nz elements");
        if (born && (numel(G_in) == 0 || G_in->dims[0] * G_in->dims[1] != P || dim(G_in, 2) < g.nt))
            raise(FD2D_EINVAL, "Born mode needs G_in with nt slices of nx x nz (run_forward1_green.m:141)");
        sid = node_ids(src, g, "src");
        rid = node_ids(rec, g, "rec");
        nr = (int64_t)rid.size();
        const int64_t n_fw = green_n_fw(g);
        check_out(G_fft, g.nx, g.nz, g.K, "G_fft");
        if (mode != 0) check_out(G_out, g.nx, g.nz, n_fw, "G_out");
        check_out(seis, nr, g.nt, 1, "seismograms");
        Gf = (G_fft && G_fft->data && g.K > 0) ? (double2*)G_fft->data : nullptr;
        Gout = (mode != 0 && G_out && G_out->data) ? (float*)G_out->data : nullptr;
        sg = (seis && seis->data) ? (double*)seis->data : nullptr;
        records = sg && nr;
        nsteps = g.nt;

        S.reset(new Solver(cfg, g, st, false));
        S->material((const double*)mu->data, (const double*)rho->data);
        if (!born) S->set_sources(sid);
        S->set_receivers(rid);
        if (born) {
            dtdmu.alloc(g.P);
            launch_scale_to_float((const double*)dmu->data, dtdmu.p, g.P, g.dt, st);
            ndtdmu.alloc(g.P);
            launch_scale_to_float((const double*)dmu->data, ndtdmu.p, g.P, -g.dt, st);
        }
        if (records) HIPCHK(hipMemsetAsync(sg, 0, sizeof(double) * (size_t)nr * g.nt, st));
        // sampled steps of :176 and their times t(n) = (n-1) dt (fft_coeff, :81-90)
        std::vector<double> ts;
        for (int n = 1; n <= g.nt; ++n)
            if ((n + g.nt - 1) % g.s == 0) ts.push_back((n - 1) * g.dt);
        // G_fft = conj(sum v * (2pi)^-1/2 e^{-iwt} dt)  (:178-180, :196-198)
        if (Gf) dft.init(Gf, g.P, cfg, ts, -1.0, 1.0, Gf, kInvSqrt2Pi * g.dt, 1, st);
    }
    bool sampled(int n) const { return Gf != nullptr && (n + g.nt - 1) % g.s == 0; }
    void plan(int n, int sub) override
    {
        const Solver::Slots sl = S->slots(sub);
        if (mode != 0 && (n + g.nt - 1) % g.f == 0) {
            if (Gout) *sl.snap = Gout + (size_t)i_fw_out * g.P;
            ++i_fw_out;
        }
        if (born) {
            // sigma += dt dmu dx_v(G_in(:,:,n)) before step n's divergence (:140-144): for n = 1 a
            // launch of its own, for the later steps inside the previous step's stress update, which
            // is the last change to sigma before it
            if (n == 1) S->born_green(dtdmu.p, (const float*)G_in->data);
            if (n < g.nt) {
                S->a.born_dmu = ndtdmu.p;
                S->a.born_in0 = (const float*)G_in->data + (size_t)n * g.P;
            }
        } else if (!sid.empty()) {
            S->point_sources(sub, nullptr, 1.0e9f);  // stf = 1e9 every step (:64)
        }
        if (sampled(n)) *sl.dsamp = dft.next_slot();
        if (records) S->receivers(sub, sg + (size_t)nr * (n - 1));
    }
    bool pairable(int n) const override { return !(sampled(n) && sampled(n + 1)); }
    void after(int n, int k) override
    {
        for (int q = 0; q < k; ++q)
            if (sampled(n + q)) dft.commit();
    }
    void finish() override
    {
        if (Gf) dft.finish();
    }
};

// =============================================================================================
// run_forward2_correlation (code/run_forward2_correlation.m)
// =============================================================================================
struct CorrelationJob : Job {
    double mode;
    const fd2d_carray* C_in;
    std::vector<int2> rid;
    int64_t nr = 0, i_fw_out = 0;
    int N = 0;
    float* Cout = nullptr;
    double* sg = nullptr;
    DBuf<double> I_d, ss_d;
    Synthesiser syn;
    DBuf<float> dmu_f;

    CorrelationJob(const fd2d_config* c, const fd2d_carray* mu, const fd2d_carray* rho, const fd2d_carray* G,
                   const fd2d_carray* spectrum, const fd2d_carray* sd, const fd2d_carray* rec, double mode_,
                   const fd2d_carray* dmu, const fd2d_carray* C_in_, fd2d_marray* seis, fd2d_marray* C_out,
                   hipStream_t s)
        : mode(mode_), C_in(C_in_)
    {
        cfg = c;
        st = s;
        solver = FD2D_CORRELATION;
        g = geometry(cfg);
        const int64_t P = (int64_t)g.P;
        N = 2 * g.nt - 1;  // :68-69
        const int nb = cfg->n_basis_fct, nb_cols = nb == 0 ? 1 : nb;
        if (numel(mu) != P || numel(rho) != P)
            raise(FD2D_EINVAL, "mu and rho must have nx*nz = %lld elements", (long long)P);
        if (numel(G) != P * g.K || g.K == 0) raise(FD2D_EINVAL, "G_fft must be nx x nz x n_sample");
        const int64_t n_src = dim(spectrum, 1);
        if (dim(spectrum, 0) != g.K || n_src < 1 || dim(spectrum, 2) != 1)
            raise(FD2D_EINVAL, "spectrum must be n_sample x n_noise_sources");
        if (nb == 0 && n_src != 1)
            raise(FD2D_EINVAL, "with n_basis_fct = 0 the reference accepts one noise source (run_forward2_correlation.m:171)");
        if (numel(sd) != P * (nb == 0 ? n_src : nb))
            raise(FD2D_EINVAL, "source_dist must be nx x nz x %d", nb == 0 ? (int)n_src : nb);
        born = numel(dmu) != 0;
        if (born && (numel(dmu) != P || numel(C_in) == 0 || C_in->dims[0] * C_in->dims[1] != P || dim(C_in, 2) < N))
            raise(FD2D_EINVAL, "Born mode needs dmu (nx x nz) and C_in with 2nt-1 slices (run_forward2_correlation.m:236)");
        rid = node_ids(rec, g, "rec");
        nr = (int64_t)rid.size();
        const int64_t n_fw = N / g.f;
        if (mode != 0) check_out(C_out, g.nx, g.nz, n_fw, "C_out");
        check_out(seis, nr, N, 1, "seismograms");
        Cout = (mode != 0 && C_out && C_out->data) ? (float*)C_out->data : nullptr;
        sg = (seis && seis->data) ? (double*)seis->data : nullptr;
        records = sg && nr;
        nsteps = N;

        S.reset(new Solver(cfg, g, st, true, displacement_form(g, born)));
        S->material((const double*)mu->data, (const double*)rho->data);
        S->set_receivers(rid);
        if (records) HIPCHK(hipMemsetAsync(sg, 0, sizeof(double) * (size_t)nr * N, st));
        // source spectrum A = (sum_ns spectrum) * distribution .* G_fft (:168-173, :201-205)
        const std::vector<double> spec = to_host<double>(spectrum->data, (size_t)g.K * n_src);
        std::vector<double> spec_sum((size_t)g.K, 0.0);
        for (int k = 0; k < g.K; ++k)
            for (int64_t q = 0; q < n_src; ++q) spec_sum[k] += spec[(size_t)k + (size_t)g.K * q];
        upload(I_d, interpol_matrix(g.K, nb), st);
        upload(ss_d, spec_sum, st);
        launch_source_spectrum(syn.spectrum_rows(g.P, g.K), (const double2*)G->data, 0, (const double*)sd->data,
                               nb_cols, I_d.p, ss_d.p, g.P, g.K, syn.C_pad, st);
        // ifft_coeff of :73-85 at t(n) = (n - nt) dt for n = s, 2s, ...
        const int n_ftc = N / g.s;
        std::vector<double> ts;
        for (int q = 1; q <= n_ftc; ++q) ts.push_back((double)(q * g.s - g.nt) * g.dt);
        syn.init(cfg, ts, kInvSqrt2Pi * g.dw, st);
        if (born) {
            dmu_f.alloc(g.P);
            launch_scale_to_float((const double*)dmu->data, dmu_f.p, g.P, 1.0, st);
        }
    }
    void plan(int n, int sub) override
    {
        const Solver::Slots sl = S->slots(sub);
        if (mode != 0 && n % g.f == 0) {  // :182-186
            if (Cout) *sl.snap = Cout + (size_t)i_fw_out * g.P;
            ++i_fw_out;
        }
        if (n % g.s == 0) *sl.syn = syn.slice(n / g.s - 1);  // :195-216
        if (born && n != N) {                                 // :232-238
            S->a.born_dmu = dmu_f.p;
            S->a.born_in0 = (const float*)C_in->data + (size_t)(n - 1) * g.P;
            S->a.born_in1 = (const float*)C_in->data + (size_t)n * g.P;
        }
        if (records) S->receivers(sub, sg + (size_t)nr * (n - 1));
    }
    bool pairable(int n) const override { return !(n % g.s == 0 && (n + 1) % g.s == 0); }
    void after(int, int) override {}
    void finish() override {}
};

// =============================================================================================
// run_noise_adjoint (code/run_noise_adjoint.m)
// =============================================================================================
struct AdjointJob : Job {
    int imode = 0, N = 0, nb_cols = 1;
    const fd2d_carray *u_fwd, *dmu, *u_in, *G;
    bool save_mode = false, stf_point = false, dist_src = false, src_kernel = false, need_dft = false;
    std::vector<int2> sid, rid;
    std::vector<double> spec;
    int64_t nr = 0, n_ufwd = 0, i_fw_in = 0, i_fw_out = 0;
    double* Kd = nullptr;
    double2* sfft = nullptr;
    float* uo = nullptr;
    double* sg = nullptr;
    DBuf<float> amp;
    DBuf<double> I_d, ss_d, sp_d;
    DBuf<double2> acc_tmp;
    Synthesiser syn;
    DftAccumulator dft;
    DBuf<float> dmu_f;
    const float* kfwd[2] = {nullptr, nullptr};
    // displacement form: the forward snapshot of the product of the last planned step, which the
    // next launch computes from its input displacement (or finish(), after the last step)
    const float* deferred = nullptr;

    AdjointJob(const fd2d_config* c, const fd2d_carray* mu, const fd2d_carray* rho, const fd2d_carray* u_fwd_,
               const fd2d_carray* stf, const fd2d_carray* src, const fd2d_carray* rec, const fd2d_carray* spectrum,
               const fd2d_carray* sd, const fd2d_carray* G_, double mode, const fd2d_carray* dmu_,
               const fd2d_carray* u_in_, fd2d_marray* K_out, fd2d_marray* stf_fft, fd2d_marray* u_out,
               fd2d_marray* seis, hipStream_t s)
        : u_fwd(u_fwd_), dmu(dmu_), u_in(u_in_), G(G_)
    {
        cfg = c;
        st = s;
        solver = FD2D_ADJOINT;
        g = geometry(cfg);
        const int64_t P = (int64_t)g.P;
        N = 2 * g.nt - 1;  // :47-49
        const int nb = cfg->n_basis_fct;
        nb_cols = nb == 0 ? 1 : nb;
        if (numel(mu) != P || numel(rho) != P)
            raise(FD2D_EINVAL, "mu and rho must have nx*nz = %lld elements", (long long)P);
        imode = (int)mode;
        if ((double)imode != mode || !(imode == 0 || imode == 1 || (imode >= 10 && imode <= 13)))
            raise(FD2D_EINVAL, "mode must be 0, 1, 10, 11, 12 or 13 (run_noise_adjoint.m:16-22)");
        const bool fft_mode = imode == 1 || imode == 12 || imode == 13;
        save_mode = imode == 11 || imode == 13;

        const int64_t n_stf = numel(stf);
        stf_point = n_stf != 0 && stf->dims[2] == 1;            // :249
        const bool stf_dist = n_stf != 0 && stf->dims[2] != 1;  // :255
        if (stf_dist && fft_mode)
            raise(FD2D_EUNSUPPORTED, "a distributed stf with mode %d advances the reference's DFT counter twice per "
                                     "sample (run_noise_adjoint.m:265,343)", imode);
        if (stf_dist && (stf->dims[0] * stf->dims[1] != P || stf->dims[2] != g.K))
            raise(FD2D_EINVAL, "a distributed stf must be nx x nz x n_sample");
        nsteps = ((imode == 0 || imode == 1) && stf_dist) ? g.nt : N;  // :165-167

        sid = node_ids(src, g, "src");
        rid = node_ids(rec, g, "rec");
        nr = (int64_t)rid.size();
        if (stf_point && (stf->dims[0] < (int64_t)sid.size() || stf->dims[1] < nsteps))
            raise(FD2D_EINVAL, "a point stf must be n_src x (2nt-1)");

        const bool have_sd = numel(sd) != 0;  // :112-123
        const int64_t n_src = have_sd ? dim(spectrum, 1) : 0;
        if (numel(spectrum)) spec = to_host<double>(spectrum->data, (size_t)numel(spectrum));
        if (have_sd) {
            if (dim(spectrum, 0) != g.K) raise(FD2D_EINVAL, "spectrum must have n_sample rows");
            if (nb == 0 && n_src != 1) raise(FD2D_EINVAL, "with n_basis_fct = 0 the reference accepts one noise source");
            if (numel(sd) != P * (nb == 0 ? n_src : nb)) raise(FD2D_EINVAL, "source_dist size");
        }
        src_kernel = numel(G) != 0 && dim(G, 2) == g.K && G->dims[0] * G->dims[1] == P && g.K > 0;  // :302
        if (src_kernel && (dim(spectrum, 0) != g.K || dim(spectrum, 1) != 1))
            raise(FD2D_EINVAL, "the source kernel needs an n_sample x 1 spectrum (run_noise_adjoint.m:312)");
        n_ufwd = numel(u_fwd) ? dim(u_fwd, 2) : 0;  // :324
        if (n_ufwd && u_fwd->dims[0] * u_fwd->dims[1] != P) raise(FD2D_EINVAL, "u_fwd slices must have nx*nz elements");
        const int64_t n_uin = numel(u_in) ? dim(u_in, 2) : 0;
        born = numel(dmu) != 0 && n_uin == N;  // :285
        if (born && (numel(dmu) != P || u_in->dims[0] * u_in->dims[1] != P)) raise(FD2D_EINVAL, "dmu / u_in sizes");

        const int64_t n_fw = N / g.f;
        check_out(K_out, g.nx, g.nz, nb_cols + 1, "K");
        if (fft_mode) check_out(stf_fft, g.nx, g.nz, g.K, "stf_fft");
        if (save_mode) check_out(u_out, g.nx, g.nz, n_fw, "u_out");
        check_out(seis, nr, N, 1, "seismograms");
        Kd = (K_out && K_out->data) ? (double*)K_out->data : nullptr;
        sfft = (fft_mode && stf_fft && stf_fft->data && g.K > 0) ? (double2*)stf_fft->data : nullptr;
        uo = (save_mode && u_out && u_out->data) ? (float*)u_out->data : nullptr;
        sg = (seis && seis->data) ? (double*)seis->data : nullptr;
        records = sg && nr;

        S.reset(new Solver(cfg, g, st, true, displacement_form(g, born)));
        S->material((const double*)mu->data, (const double*)rho->data);
        if (Kd) {
            S->a.kacc = Kd + (size_t)nb_cols * g.P;  // K_mu (:359-367)
            S->a.kscale = (double)g.f;
        }
        if (stf_point) S->set_sources(sid);
        S->set_receivers(rid);
        if (records) HIPCHK(hipMemsetAsync(sg, 0, sizeof(double) * (size_t)nr * N, st));
        if (Kd) HIPCHK(hipMemsetAsync(Kd, 0, sizeof(double) * g.P * (size_t)(nb_cols + 1), st));

        // point adjoint sources: real part of stf(i, n) (:249-253)
        if (stf_point && !sid.empty()) {
            amp.alloc(sid.size() * (size_t)nsteps);
            launch_stf_real(amp.p, (const double2*)stf->data, stf->dims[0], (int)sid.size(), nsteps, st);
        }
        // distributed adjoint source: A = (sum_ns spectrum) * distribution .* conj(stf) (:255-268)
        const std::vector<double> I = interpol_matrix(g.K, nb);
        upload(I_d, I, st);
        dist_src = stf_dist && have_sd;
        if (dist_src) {
            std::vector<double> spec_sum((size_t)g.K, 0.0);
            for (int k = 0; k < g.K; ++k)
                for (int64_t q = 0; q < n_src; ++q) spec_sum[k] += spec[(size_t)k + (size_t)g.K * q];
            upload(ss_d, spec_sum, st);
            launch_source_spectrum(syn.spectrum_rows(g.P, g.K), (const double2*)stf->data, 1, (const double*)sd->data,
                                   nb_cols, I_d.p, ss_d.p, g.P, g.K, syn.C_pad, st);
            const int n_syn = nsteps / g.s;
            std::vector<double> ts;
            for (int q = 1; q <= n_syn; ++q) ts.push_back((double)(q * g.s - g.nt) * g.dt);
            syn.init(cfg, ts, kInvSqrt2Pi * g.dw, st);
        }
        // DFT of u at the sampled steps: stf_fft (:335-346) and the source kernel (:302-320) share it
        need_dft = (sfft != nullptr) || (src_kernel && Kd);
        if (need_dft) {
            double2* accp = sfft;
            if (!accp) {
                acc_tmp.alloc(g.P * (size_t)g.K);
                accp = acc_tmp.p;
            }
            std::vector<double> ts;
            for (int n = g.s; n <= nsteps; n += g.s) ts.push_back((double)(n - g.nt) * g.dt);
            // with a source kernel, the series ends as the plain sums (read by the kernel, then
            // scaled into stf_fft); else directly as stf_fft = sum u * (2pi)^-1/2 e^{-iwt} dt
            if (src_kernel && Kd) dft.init(accp, g.P, cfg, ts, -1.0, 1.0, accp, 1.0, 0, st);
            else dft.init(accp, g.P, cfg, ts, -1.0, 1.0, accp, kInvSqrt2Pi * g.dt, 0, st);
        }
        if (born) {
            dmu_f.alloc(g.P);
            launch_scale_to_float((const double*)dmu->data, dmu_f.p, g.P, 1.0, st);
        }
    }
    bool kernel_step(int n) const { return Kd && n_ufwd && n_ufwd >= i_fw_in + 1 && n % g.f == 0; }  // :324
    bool product_at(int n) const override { return kernel_step(n); }
    bool sampled(int n) const { return need_dft && n % g.s == 0; }
    bool synth(int n) const { return !(stf_point && !sid.empty()) && dist_src && n % g.s == 0; }
    void plan(int n, int sub) override
    {
        const Solver::Slots sl = S->slots(sub);
        if (save_mode && n % g.f == 0) {  // :237-241
            if (uo) *sl.snap = uo + (size_t)i_fw_out * g.P;
            ++i_fw_out;
        }
        if (stf_point && !sid.empty()) {
            S->point_sources(sub, amp.p + (size_t)(n - 1) * sid.size(), 0.f);
        } else if (synth(n)) {
            *sl.syn = syn.slice(n / g.s - 1);
        }
        if (born && n != N) {  // :285-291
            S->a.born_dmu = dmu_f.p;
            S->a.born_in0 = (const float*)u_in->data + (size_t)(n - 1) * g.P;
            S->a.born_in1 = (const float*)u_in->data + (size_t)n * g.P;
        }
        kfwd[sub] = nullptr;
        if (S->uform) S->a.kfwd = deferred;  // the previous step's product, in this launch
        if (kernel_step(n)) {  // :324-331
            kfwd[sub] = (const float*)u_fwd->data + (size_t)(n_ufwd - 1 - i_fw_in) * g.P;
            ++i_fw_in;
        }
        if (S->uform) {
            deferred = kfwd[sub];
            kfwd[sub] = nullptr;
        }
        if (sampled(n)) *sl.dsamp = dft.next_slot();
        if (records) S->receivers(sub, sg + (size_t)nr * (n - 1));
    }
    // A pair needs the structure-kernel product, if any, after its second step (the displacement
    // of its first step is not kept), at most one DFT sample and at most one synthesised source.
    bool pairable(int n) const override
    {
        return !kernel_step(n) && !(sampled(n) && sampled(n + 1)) && !(synth(n) && synth(n + 1));
    }
    void fuse_products() override
    {
        if (S->uform) return;  // plan() has set the deferred product
        S->a.kfwd = kfwd[0];
        kfwd[0] = nullptr;
    }
    void after(int n, int k) override
    {
        // a product left for after the launch: the second step of a two-step launch
        if (kfwd[k - 1]) S->kmu(kfwd[k - 1], Kd + (size_t)nb_cols * g.P, (double)g.f);
        for (int q = 0; q < k; ++q)
            if (sampled(n + q)) dft.commit();
    }
    void finish() override
    {
        if (deferred) S->kmu(deferred, Kd + (size_t)nb_cols * g.P, (double)g.f);  // the last step's product
        deferred = nullptr;
        if (!need_dft) return;
        dft.finish();
        if (src_kernel && Kd) {
            // K_s(:,:,b) = sum_n Re(M_b(n) u(n)) = (2pi)^-1/2 dw sum_k I(k,b) spec_k Re(conj(G_k D_k)),
            // D_k = sum_n u(n) e^{-i w_k t_n} (the finished sums in dft.acc)
            upload(sp_d, std::vector<double>(spec.begin(), spec.begin() + g.K), st);
            launch_source_kernel(Kd, (const double2*)G->data, dft.acc, nullptr, I_d.p, sp_d.p, g.P, g.K, nb_cols,
                                 kInvSqrt2Pi * g.dw, st);
            if (sfft) launch_finalize_dft(sfft, sfft, nullptr, g.P * (size_t)g.K, kInvSqrt2Pi * g.dt, 0, st);
        }
    }
};

void green_device(const fd2d_config* cfg, const fd2d_carray* mu, const fd2d_carray* rho, const fd2d_carray* src,
                  const fd2d_carray* rec, double mode, const fd2d_carray* dmu, const fd2d_carray* G_in,
                  fd2d_marray* G_fft, fd2d_marray* G_out, fd2d_marray* seis, hipStream_t st)
{
    StreamScope scope(st);
    GreenJob j(cfg, mu, rho, src, rec, mode, dmu, G_in, G_fft, G_out, seis, st);
    run_job(j);
}

void correlation_device(const fd2d_config* cfg, const fd2d_carray* mu, const fd2d_carray* rho, const fd2d_carray* G,
                        const fd2d_carray* spectrum, const fd2d_carray* sd, const fd2d_carray* rec, double mode,
                        const fd2d_carray* dmu, const fd2d_carray* C_in, fd2d_marray* seis, fd2d_marray* C_out,
                        hipStream_t st)
{
    StreamScope scope(st);
    CorrelationJob j(cfg, mu, rho, G, spectrum, sd, rec, mode, dmu, C_in, seis, C_out, st);
    run_job(j);
}

void adjoint_device(const fd2d_config* cfg, const fd2d_carray* mu, const fd2d_carray* rho, const fd2d_carray* u_fwd,
                    const fd2d_carray* stf, const fd2d_carray* src, const fd2d_carray* rec,
                    const fd2d_carray* spectrum, const fd2d_carray* sd, const fd2d_carray* G, double mode,
                    const fd2d_carray* dmu, const fd2d_carray* u_in, fd2d_marray* K_out, fd2d_marray* stf_fft,
                    fd2d_marray* u_out, fd2d_marray* seis, hipStream_t st)
{
    StreamScope scope(st);
    AdjointJob j(cfg, mu, rho, u_fwd, stf, src, rec, spectrum, sd, G, mode, dmu, u_in, K_out, stf_fft, u_out, seis, st);
    run_job(j);
}

// =============================================================================================
// Batch engine: many solves advanced together. Every round advances each active solve by one
// time step; solves with the same FD order and field set are stepped by one grouped launch
// (k_tile_group) over the tiles of all of them, the rest one launch each. A solve's results are
// those of its own entry point bit for bit (the same tile kernel body runs on the same data).
// =============================================================================================
struct Entry {
    fd2d_config cfg{};
    std::vector<double> w;
    fd2d_carray in[11]{};
    bool has_in[11]{};
    fd2d_marray out[4]{};
    bool has_out[4]{};
    std::unique_ptr<Job> job;
    int64_t ticket = 0;
    int slot = -1;
    StepArgs stat[2]{};             // host copy of the static arguments written to the slot
    hipEvent_t done_ev = nullptr;   // recorded after the solve's last work
    hipEvent_t chk_ev = nullptr;    // recorded after the copy of its last finiteness check
    int lane = 0;                   // the engine stream its work runs on
    bool chk_pending = false;       // a check is queued and chk_ev not yet seen complete
    int next_check = 0;             // steps issued at which the next periodic check is queued
    ~Entry()
    {
        if (done_ev) (void)hipEventDestroy(done_ev);
        if (chk_ev) (void)hipEventDestroy(chk_ev);
    }
    const fd2d_carray* cin(int q) const { return has_in[q] ? &in[q] : nullptr; }
    fd2d_marray* mout(int q) { return has_out[q] ? &out[q] : nullptr; }
};

enum { kMu, kRho, kSrc, kRec, kGfft, kSpec, kSd, kStf, kUfwd, kDmu, kBornIn };

void check_solve(const fd2d_solve* s)
{
    if (!s || !s->cfg) raise(FD2D_EINVAL, "solve and its cfg must not be NULL");
    if (s->solver < FD2D_GREEN || s->solver > FD2D_ADJOINT)
        raise(FD2D_EINVAL, "solver must be FD2D_GREEN, FD2D_CORRELATION or FD2D_ADJOINT");
    if (s->cfg->n_sample > 0 && !s->cfg->w_sample) raise(FD2D_EINVAL, "cfg->w_sample is NULL");
}

std::unique_ptr<Entry> make_entry(const fd2d_solve* s, hipStream_t st)
{
    check_solve(s);
    std::unique_ptr<Entry> e(new Entry);
    e->cfg = *s->cfg;
    if (e->cfg.n_sample > 0) e->w.assign(s->cfg->w_sample, s->cfg->w_sample + e->cfg.n_sample);
    e->cfg.w_sample = e->w.data();
    const fd2d_carray* src_in[11] = {s->mu, s->rho, s->src, s->rec, s->G_fft, s->spectrum, s->source_dist,
                                     s->stf, s->u_fwd, s->dmu, s->born_in};
    for (int q = 0; q < 11; ++q)
        if (src_in[q]) {
            e->in[q] = *src_in[q];
            e->has_in[q] = true;
        }
    for (int q = 0; q < 4; ++q)
        if (s->out[q]) {
            e->out[q] = *s->out[q];
            e->has_out[q] = true;
        }
    const fd2d_config* c = &e->cfg;
    Entry& E = *e;
    switch (s->solver) {
    case FD2D_GREEN:
        E.job.reset(new GreenJob(c, E.cin(kMu), E.cin(kRho), E.cin(kSrc), E.cin(kRec), s->mode, E.cin(kDmu),
                                 E.cin(kBornIn), E.mout(0), E.mout(1), E.mout(2), st));
        break;
    case FD2D_CORRELATION:
        E.job.reset(new CorrelationJob(c, E.cin(kMu), E.cin(kRho), E.cin(kGfft), E.cin(kSpec), E.cin(kSd),
                                       E.cin(kRec), s->mode, E.cin(kDmu), E.cin(kBornIn), E.mout(0), E.mout(1), st));
        break;
    default:
        E.job.reset(new AdjointJob(c, E.cin(kMu), E.cin(kRho), E.cin(kUfwd), E.cin(kStf), E.cin(kSrc), E.cin(kRec),
                                   E.cin(kSpec), E.cin(kSd), E.cin(kGfft), s->mode, E.cin(kDmu), E.cin(kBornIn),
                                   E.mout(0), E.mout(1), E.mout(2), E.mout(3), st));
        break;
    }
    return e;
}

constexpr int kEngineSlots = 1024;  // solves an engine holds at once
// Tiles from which a solve runs alone in the engine: 512 tiles of 512 threads already occupy every
// CU with two workgroups. Measured at 1200 x 1200 (722 tiles, 8 solves in flight): a grouped launch
// takes 14.2 us per solve-step against 13.1 us for the solve's own launch, and stepping the 8 solves
// in turn with their own launches is slower still.
constexpr int kSoloTiles = 512;
// MI355X's last-level (Infinity) cache; large solves are grouped while their step working sets fit
constexpr double kLastLevelCacheBytes = 256.0 * (1 << 20);
// fields and coefficients one step reads and writes: 32 B per node without the displacement field
// (Green runs), 40 B with it
// (the displacement form moves 32 B: v and the FP64 u)
double step_bytes(const Job& j) { return (double)j.g.P * (field_kind(j) == 1 ? 40.0 : 32.0); }

// Steps between the periodic finiteness checks of an engine's solve (each reads the solve's field
// pool once: 1/kCheckEvery of a step's traffic).
constexpr int kCheckEvery = 512;

// Streams ("lanes") of an engine: each solve runs on the lane of its FD order and field set, so the
// launches of different classes -- which cannot share a launch -- run side by side on the GPU
// (small grids fill a fraction of the CUs each); the solves of one grouped launch share a lane.
constexpr int kLanes = 4;
int lane_of(int order, int solver)
{
    const int R = order / 2;
    return (solver == FD2D_GREEN ? 0 : 1) + 2 * ((R >> 1) & 1);  // orders 4 and 8 (and 2 and 6) apart
}

struct Engine {
    hipStream_t st = nullptr;          // the caller's stream: inputs are ready on it at submission
    hipStream_t lanes[kLanes] = {};
    hipEvent_t submit_ev = nullptr;    // recorded on st at each submission, waited for by the lane
    std::vector<std::unique_ptr<Entry>> act;       // running
    std::vector<std::unique_ptr<Entry>> draining;  // finished, cancelled or failed; their last work still queued
    // Solves that were cancelled or failed, with the steps they had issued. Complete solves keep no
    // record: a ticket below next_ticket that is neither running, draining nor here is complete.
    struct Closed {
        int state;
        int64_t issued, total;
    };
    std::map<int64_t, Closed> closed;
    int64_t next_ticket = 1;
    // finiteness checks: per-slot flags on the device, copied into pinned host memory before the
    // check's event; the partials are the scratch of the check kernel (one per lane: one check at a
    // time per stream)
    DBuf<int> chk_partial[kLanes], chk_flag_d;
    int* chk_flag_h = nullptr;
    bool dirty = true;
    // Static StepArgs of every solve (both parities) at a fixed slot, written once at submission:
    // grouped launches index it with the slot each JobStep carries.
    DBuf<StepArgs> slots_d;
    std::vector<int> free_slots;
    struct Group {
        int R = 0;
        int kind = 0;  // field_kind of its members
        std::vector<Entry*> members;
    };
    std::vector<Group> groups;
    std::vector<Entry*> singles;

    explicit Engine(hipStream_t s) : st(s)
    {
        StreamScope scope(st);
        for (int q = 0; q < kLanes; ++q) HIPCHK(hipStreamCreateWithFlags(&lanes[q], hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&submit_ev, hipEventDisableTiming));
        slots_d.alloc((size_t)2 * kEngineSlots);
        for (int q = 0; q < kLanes; ++q) chk_partial[q].alloc(kCheckBlocks);
        chk_flag_d.alloc(kEngineSlots);
        HIPCHK(hipHostMalloc((void**)&chk_flag_h, sizeof(int) * kEngineSlots, hipHostMallocDefault));
        std::fill(chk_flag_h, chk_flag_h + kEngineSlots, 0);
        for (int q = kEngineSlots - 1; q >= 0; --q) free_slots.push_back(q);
    }
    // Queue a finiteness check of e's field pool; its flag reaches chk_flag_h[slot] by chk_ev.
    void queue_check(Entry& e)
    {
        const Solver& S = *e.job->S;
        const hipStream_t ls = e.job->st;
        launch_nonfinite(S.check_ptr(), S.check_count(), chk_partial[e.lane].p, chk_flag_d.p + e.slot, ls);
        HIPCHK(hipMemcpyAsync(chk_flag_h + e.slot, chk_flag_d.p + e.slot, sizeof(int), hipMemcpyDeviceToHost, ls));
        if (!e.chk_ev) HIPCHK(hipEventCreateWithFlags(&e.chk_ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(e.chk_ev, ls));
        e.chk_pending = true;
    }
    // Take a running solve out (cancelled or failed): no further steps; its queued work drains.
    void close(std::vector<std::unique_ptr<Entry>>::iterator it, int state)
    {
        Entry& e = **it;
        closed[e.ticket] = Closed{state, (int64_t)e.job->next - 1, (int64_t)e.job->nsteps};
        if (!e.done_ev) HIPCHK(hipEventCreateWithFlags(&e.done_ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(e.done_ev, e.job->st));
        draining.push_back(std::move(*it));
        act.erase(it);
        dirty = true;
    }
    // Periodic checks: fail the solves whose last check found a NaN or Inf, and queue the next
    // check of every solve that has issued kCheckEvery more steps. A failed solve's fields are its
    // own, so the solves batched with it are not affected.
    void poll_checks()
    {
        for (size_t q = 0; q < act.size();) {
            Entry& e = *act[q];
            if (e.chk_pending) {
                const hipError_t r = hipEventQuery(e.chk_ev);
                if (r != hipErrorNotReady) {
                    HIPCHK(r);
                    e.chk_pending = false;
                    if (chk_flag_h[e.slot]) {
                        close(act.begin() + q, FD2D_SOLVE_FAILED);
                        continue;
                    }
                }
            }
            if (!e.chk_pending && e.job->next - 1 >= e.next_check) {
                queue_check(e);
                e.next_check += kCheckEvery;
            }
            ++q;
        }
    }
    // Re-partition the running solves after the set changed (host work only).
    void rebuild()
    {
        groups.clear();
        singles.clear();
        // class of a solve: its lane, FD order and field set. A grouped launch runs on its members'
        // lane, which must be the lane of every member: their earlier steps, and the events that
        // free their buffers, are queued there.
        std::vector<std::pair<std::array<int, 3>, Entry*>> cls;
        std::vector<Entry*> big;  // the oldest large solves of one class
        double big_bytes = 0.0;
        for (auto& e : act) {
            Job& j = *e->job;
            if (j.next > j.nsteps) continue;
            // Solves that fill the chip on their own step together, the oldest first, only as many
            // as the last-level cache holds the per-step working sets of: more of them evict each
            // other's fields from it; up to that size, one launch over their tiles fills the last
            // round of workgroups that one solve leaves part-empty.
            if (j.S->n_tiles >= kSoloTiles) {
                const double b = step_bytes(j);
                if (big.empty() || (big_bytes + b <= kLastLevelCacheBytes && big[0]->lane == e->lane &&
                                    big[0]->job->g.R == j.g.R &&
                                    field_kind(*big[0]->job) == field_kind(j) && !j.born && !big[0]->job->born)) {
                    big.push_back(e.get());
                    big_bytes += b;
                }
                continue;
            }
            cls.push_back({{e->lane, j.g.R, field_kind(j)}, e.get()});
        }
        if (big.size() == 1) {
            singles.push_back(big[0]);
        } else if (big.size() > 1) {
            Group gr;
            gr.R = big[0]->job->g.R;
            gr.kind = field_kind(*big[0]->job);
            gr.members = big;
            groups.push_back(std::move(gr));
        }
        std::stable_sort(cls.begin(), cls.end(),
                         [](const auto& x, const auto& y) { return x.first < y.first; });
        // a class of n solves steps in ceil(n / kPairGroupMax) launches of near-equal size (64 solves:
        // two launches of 32, not 36 + 24 + 4, whose smallest takes a third of the largest's time)
        static_assert(kPairGroupMax <= kGroupMax, "a group must fit either grouped launch");
        for (size_t q = 0; q < cls.size();) {
            size_t r = q;
            while (r < cls.size() && cls[r].first == cls[q].first) ++r;
            const size_t n = r - q, parts = (n + kPairGroupMax - 1) / kPairGroupMax;
            for (size_t p = 0; p < parts; ++p) {
                const size_t b = q + n * p / parts, e = q + n * (p + 1) / parts;
                if (e - b == 1) {
                    singles.push_back(cls[b].second);
                    continue;
                }
                Group gr;
                gr.R = cls[q].first[1];
                gr.kind = cls[q].first[2];
                for (size_t k = b; k < e; ++k) gr.members.push_back(cls[k].second);
                groups.push_back(std::move(gr));
            }
            q = r;
        }
        dirty = false;
    }
    // A solve with its own launch: one step, or two when its launch shape (shape_of) is two steps
    // and the pair is allowed.
    void step_single(Job& j)
    {
        StreamScope scope(j.st);
        const int k = (shape_of(j) == 2 && !j.born && j.next + 1 <= j.nsteps && j.pairable(j.next)) ? 2 : 1;
        j.S->clear_step();
        for (int q = 0; q < k; ++q) j.plan(j.next + q, q);
        if (k == 1) {
            j.fuse_products();
            j.S->step();
        } else {
            j.S->step_pair();
        }
        j.after(j.next, k);
        j.next += k;
    }
    // Steps per launch the engine uses for solve j: the tuned shape of its grid when an earlier solve
    // has timed it (or the fd2d_set_steps_per_launch setting), else two steps at FD order 2 / 4 and
    // one at 6 / 8 (measured: two-step launches are faster at order 4 on every grid tried, equal at
    // order 8 on the large benchmark grid).
    int shape_of(const Job& j) const
    {
        const int mode = g_steps_per_launch.load();
        if (j.S->uform) return 1;  // the displacement form has single-step launches only
        if (mode == 1 || mode == 2) return mode;
        const int t = tuned_steps(j);
        return t ? t : (j.g.R <= 2 ? 2 : 1);
    }
    // whether solve j takes steps next, next+1 in one grouped two-step launch: allowed by the pair
    // rules, no Born terms, and no structure-kernel product at the second step (which would need
    // its own launch after the pair)
    bool pairs_now(const Job& j) const
    {
        return shape_of(j) == 2 && !j.born && j.next + 1 <= j.nsteps && j.pairable(j.next) &&
               !j.product_at(j.next + 1);
    }
    void step_group(const Group& gr)
    {
        for (const Entry* e : gr.members)
            if (e->lane != gr.members[0]->lane) raise(FD2D_EINVAL, "internal: a grouped launch spans lanes");
        StreamScope scope(gr.members[0]->job->st);
        std::vector<Entry*> one, two;
        for (Entry* e : gr.members) (pairs_now(*e->job) ? two : one).push_back(e);
        if (!one.empty()) launch_singles(gr, one);
        for (size_t q = 0; q < two.size(); q += kPairGroupMax)
            launch_pairs(gr, std::vector<Entry*>(two.begin() + q, two.begin() + std::min(two.size(), q + kPairGroupMax)));
    }
    void launch_singles(const Group& gr, const std::vector<Entry*>& members)
    {
        GroupArgs ga{};
        ga.jobs = slots_d.p;
        ga.n_jobs = (int)members.size();
        int total = 0;
        for (int q = 0; q < ga.n_jobs; ++q) {
            Entry& e = *members[q];
            Job& j = *e.job;
            j.S->clear_step();
            j.plan(j.next, 0);
            j.fuse_products();
            const StepArgs& a = j.S->a;
            ga.tile_start[q] = total;
            total += j.S->n_tiles;
            ga.step[q] = JobStep{a.ps_amp, a.syn, a.snap, a.dsamp, a.seis_col, a.born_dmu, a.born_in0, a.born_in1,
                                 a.kfwd, a.ps_const, j.S->cur, e.slot};
        }
        ga.tile_start[ga.n_jobs] = total;
        ga.total_tiles = total;
        ga.tiles_per_xcd = (total + kXCD - 1) / kXCD;
        const hipStream_t ls = members[0]->job->st;  // the group's lane
        if (gr.kind == 2) launch_ugroup(gr.R, ga, ls);
        else launch_group(gr.R, gr.kind == 1, ga, ls);
        for (int q = 0; q < ga.n_jobs; ++q) {
            Job& j = *members[q]->job;
            j.S->cur ^= 1;
            j.after(j.next, 1);
            ++j.next;
        }
    }
    void launch_pairs(const Group& gr, const std::vector<Entry*>& members)
    {
        PairGroupArgs ga{};
        ga.jobs = slots_d.p;
        ga.n_jobs = (int)members.size();
        int total = 0;
        for (int q = 0; q < ga.n_jobs; ++q) {
            Entry& e = *members[q];
            Job& j = *e.job;
            j.S->clear_step();
            j.plan(j.next, 0);
            j.plan(j.next + 1, 1);
            const StepArgs& a = j.S->a;
            const PairArgs& p = j.S->pr;
            ga.tile_start[q] = total;
            total += j.S->n_tiles;
            ga.step[q] = PairStep{a.ps_amp, a.syn, a.snap, a.dsamp, a.seis_col, p.ps_amp2, p.syn2, p.snap2, p.dsamp2,
                                  p.seis_col2, a.ps_const, p.ps_const2, j.S->cur, e.slot};
        }
        ga.tile_start[ga.n_jobs] = total;
        ga.total_tiles = total;
        ga.tiles_per_xcd = (total + kXCD - 1) / kXCD;
        launch_pair_group(gr.R, gr.kind == 1, ga, members[0]->job->st);
        for (int q = 0; q < ga.n_jobs; ++q) {
            Job& j = *members[q]->job;
            j.S->cur ^= 1;  // a pair swaps the state buffers once, as one step does
            j.after(j.next, 2);
            j.next += 2;
        }
    }
    // Drop the solves whose last work has completed on the device.
    void reap()
    {
        for (auto& e : draining) {
            const hipError_t q = hipEventQuery(e->done_ev);
            if (q == hipErrorNotReady) continue;
            HIPCHK(q);
            // the final check of a finished solve was queued before done_ev
            if (chk_flag_h[e->slot] && !closed.count(e->ticket))
                closed[e->ticket] = Closed{FD2D_SOLVE_FAILED, (int64_t)e->job->nsteps, (int64_t)e->job->nsteps};
            chk_flag_h[e->slot] = 0;
            free_slots.push_back(e->slot);
            e.reset();
        }
        draining.erase(std::remove(draining.begin(), draining.end(), nullptr), draining.end());
    }
    // Finish the solves that have taken their last step: their final work and a last finiteness
    // check are queued and an event marks their completion; no host wait.
    void retire()
    {
        bool any = false;
        for (auto& e : act)
            if (e->job->next > e->job->nsteps) {
                {
                    StreamScope scope(e->job->st);
                    e->job->finish();
                }
                queue_check(*e);
                HIPCHK(hipEventCreateWithFlags(&e->done_ev, hipEventDisableTiming));
                HIPCHK(hipEventRecord(e->done_ev, e->job->st));
                draining.push_back(std::move(e));
                any = true;
            }
        if (any) {
            act.erase(std::remove(act.begin(), act.end(), nullptr), act.end());
            dirty = true;
        }
        reap();
    }
    // A class of small solves (fewer than kSoloTiles tiles each) whose running solves of one FD order
    // hold kSpillTiles or more tiles on its lane is shared with the partner lane, the lane of the
    // same solver kind for the other FD orders, while no solve of another order runs there: each new
    // solve goes to whichever of the two holds fewer of the class's tiles. The class's launches then
    // run side by side on the device, each filling the other's last, part-empty round of workgroups,
    // instead of one launch, or two in turn, that ends with one (64 Green runs at 200^2: 14.7x the
    // plain path against 9.3x with two launches of 32 in turn; 16 at 300^2: 7.9x against 5.9x with
    // one launch). Below kSpillTiles one launch is all the device holds at once anyway, and the
    // second launch only costs (8 Green runs at 200^2, 224 tiles: 5.7x split against 5.8x). A
    // partner lane busy with its own orders is left alone (a served mix that spilled onto it ran 2%
    // slower), and large solves keep their lane: they step in groups the last-level cache holds.
    static constexpr int kSpillTiles = 384;
    int spill_lane(int lane, int R, int tiles) const
    {
        if (tiles >= kSoloTiles) return lane;
        const int partner = lane ^ 2;
        int64_t n = 0, m = 0;
        bool partner_busy = false;
        for (const auto& e : act) {
            const Job& j = *e->job;
            if (j.S->n_tiles >= kSoloTiles) continue;
            if (e->lane == lane && j.g.R == R) n += j.S->n_tiles;
            if (e->lane == partner && j.g.R == R) m += j.S->n_tiles;
            if (e->lane == partner && j.g.R != R) partner_busy = true;
        }
        return !partner_busy && n >= kSpillTiles && m < n ? partner : lane;
    }
    int64_t submit(const fd2d_solve* s)
    {
        check_solve(s);
        if (free_slots.empty()) raise(FD2D_ENOMEM, "the engine holds %d solves already", kEngineSlots);
        // the solve's set-up and steps run on its lane, after the work queued on the caller's stream
        const int tiles = (s->cfg->nx + kTX - 1) / kTX * ((s->cfg->nz + kTileZ - 1) / kTileZ);
        const int lane = spill_lane(lane_of(s->cfg->order, s->solver), s->cfg->order / 2, tiles);
        const hipStream_t ls = lanes[lane];
        HIPCHK(hipEventRecord(submit_ev, st));
        HIPCHK(hipStreamWaitEvent(ls, submit_ev, 0));
        StreamScope scope(ls);
        std::unique_ptr<Entry> e = make_entry(s, ls);
        e->lane = lane;
        e->slot = free_slots.back();
        e->stat[0] = e->job->static_args(0);
        e->stat[1] = e->job->static_args(1);
        HIPCHK(hipMemcpyAsync(slots_d.p + 2 * (size_t)e->slot, e->stat, sizeof e->stat, hipMemcpyHostToDevice, ls));
        free_slots.pop_back();
        chk_flag_h[e->slot] = 0;
        e->next_check = kCheckEvery;
        e->ticket = next_ticket++;
        act.push_back(std::move(e));
        dirty = true;
        return act.back()->ticket;
    }
    void advance(int rounds)
    {
        StreamScope scope(st);
        for (int r = 0; r < rounds && !act.empty(); ++r) {
            retire();
            if (act.empty()) break;
            if (dirty) rebuild();
            for (const Group& gr : groups) step_group(gr);
            for (Entry* e : singles) step_single(*e->job);
            poll_checks();
        }
        retire();
        HIPCHK(hipGetLastError());
        if (act.empty() && !draining.empty()) {  // nothing left to overlap with: wait for the rest
            sync_lanes();
            reap();
        }
    }
    // Wait for every lane; the caller's stream then also waits for them (ordering later work on it).
    void sync_lanes()
    {
        for (int q = 0; q < kLanes; ++q) {
            HIPCHK(hipEventRecord(submit_ev, lanes[q]));
            HIPCHK(hipStreamWaitEvent(st, submit_ev, 0));
            HIPCHK(hipStreamSynchronize(lanes[q]));
        }
    }
    // solves submitted and not yet complete
    int pending() const { return (int)(act.size() + draining.size()); }
    // No longer running and its queued work complete (complete, cancelled or failed).
    bool is_done(int64_t ticket)
    {
        reap();
        for (auto& e : act)
            if (e->ticket == ticket) return false;
        for (auto& e : draining)
            if (e->ticket == ticket) return false;
        return true;
    }
    // State of a known ticket and the time steps it has issued (a host-side count: every round
    // issues one or two steps of each running solve, independent of how far the device has got).
    int state_of(int64_t ticket, int64_t* issued, int64_t* total)
    {
        reap();
        auto cl = closed.find(ticket);
        if (cl != closed.end()) {
            *issued = cl->second.issued;
            *total = cl->second.total;
            return cl->second.state;
        }
        for (auto& e : act)
            if (e->ticket == ticket) {
                *issued = e->job->next - 1;
                *total = e->job->nsteps;
                return FD2D_SOLVE_RUNNING;
            }
        for (auto& e : draining)
            if (e->ticket == ticket) {
                *issued = *total = e->job->nsteps;
                return FD2D_SOLVE_FINISHING;
            }
        *issued = *total = -1;
        return FD2D_SOLVE_COMPLETE;
    }
    // Cancel a running solve: it issues no further steps. Returns false when it is not running.
    bool cancel(int64_t ticket)
    {
        for (auto it = act.begin(); it != act.end(); ++it)
            if ((*it)->ticket == ticket) {
                close(it, FD2D_SOLVE_CANCELLED);
                reap();
                return true;
            }
        return false;
    }
    ~Engine()
    {
        for (int q = 0; q < kLanes; ++q)
            if (lanes[q] && hipStreamSynchronize(lanes[q]) != hipSuccess) (void)hipGetLastError();
        if (hipStreamSynchronize(st) != hipSuccess) (void)hipGetLastError();
        StreamScope scope(st);
        act.clear();
        draining.clear();
        slots_d.release();
        for (int q = 0; q < kLanes; ++q) chk_partial[q].release();
        chk_flag_d.release();
        if (chk_flag_h) (void)hipHostFree(chk_flag_h);
        if (submit_ev) (void)hipEventDestroy(submit_ev);
        for (int q = 0; q < kLanes; ++q)
            if (lanes[q]) {
                detach_stream(lanes[q]);
                (void)hipStreamDestroy(lanes[q]);
            }
    }
};

// Device bytes a solve allocates inside the library (fields, coefficients, lists, spectra and slabs),
// an upper bound for the sizes of its arrays; the caller's own arrays are not counted.
int64_t solve_bytes(const fd2d_solve* s)
{
    check_solve(s);
    const fd2d_config* cfg = s->cfg;
    const Geometry g = geometry(cfg);
    const PaddedLayout lay = PaddedLayout::make(g.nx, g.nz);
    const int64_t P = (int64_t)g.P, E = (int64_t)lay.elems(), K = g.K, C_pad = spec_cols(g.K);
    const bool with_u = s->solver != FD2D_GREEN;
    // field pool: two sets of v, sigma_xy, sigma_zy (, u), or of v and the FP64 u (displacement form)
    const bool uform = with_u && displacement_form(g, numel(s->dmu) != 0);
    int64_t b = E * 4 * (uform ? 3 : with_u ? 4 : 3) * 2 + E * 4 * 2 + 4 * (int64_t)(lay.pitch + lay.rows + g.nx + g.nz);
    const int64_t n_src = numel(s->src) ? dim(s->src, 0) : 0, n_rec = numel(s->rec) ? dim(s->rec, 0) : 0;
    b += (int64_t)(g.nx / kTX + 2) * (g.nz / kTileZ + 2) * 8 + 8 * 9 * (n_src + n_rec);  // tile lists
    const int N = 2 * g.nt - 1;
    auto slab = [&](int64_t n) { return (int64_t)slab_len((size_t)P, (size_t)std::max<int64_t>(n, 1)); };
    // slab, phase-table parts and the accumulators' compensation terms
    auto dft_bytes = [&](int64_t n) { return slab(n) * P * 4 + 3 * ((n + 15) / 16 * 16) * C_pad * 2 + P * K * 16; };
    auto syn_bytes = [&](int64_t n) {
        const int64_t T = slab(n), J = (T + 63) / 64 * 64, slabs = (std::max<int64_t>(n, 1) + T - 1) / T;
        return T * P * 4 + C_pad * P * 4 + 3 * slabs * C_pad * J * 2;  // slab, spectrum rows, tables
    };
    const bool born = numel(s->dmu) != 0;
    if (s->solver == FD2D_GREEN) {
        if (numel(s->out[0]) && K) b += dft_bytes(g.nt / std::max(g.s, 1) + 1);
        if (born) b += P * 8;  // dt dmu and -dt dmu
    } else if (s->solver == FD2D_CORRELATION) {
        b += syn_bytes(N / std::max(g.s, 1)) + K * (int64_t)std::max(cfg->n_basis_fct, 1) * 8 + K * 8;
        if (born) b += P * 4;
    } else {
        const bool stf_dist = numel(s->stf) && s->stf->dims[2] != 1;
        const int64_t nsteps = N;
        if (numel(s->stf) && !stf_dist) b += n_src * nsteps * 4;
        if (stf_dist && numel(s->source_dist)) b += syn_bytes(nsteps / std::max(g.s, 1));
        const int imode = (int)s->mode;
        const bool fft_mode = imode == 1 || imode == 12 || imode == 13;
        const bool src_kernel = numel(s->G_fft) != 0;
        if (fft_mode || src_kernel) {
            b += dft_bytes(nsteps / std::max(g.s, 1));
            if (!(fft_mode && numel(s->out[1]))) b += P * K * 16;
        }
        b += K * (int64_t)std::max(cfg->n_basis_fct, 1) * 8 + 2 * K * 8;
        if (born) b += P * 4;
    }
    return b;
}

// ---------------------------------------------------------------------------------------------
// Host-memory wrappers: copy every array in, run on the null stream, copy outputs back.
// ---------------------------------------------------------------------------------------------
struct Staged {
    std::vector<std::unique_ptr<DBuf<char>>> bufs;
    fd2d_carray in(const fd2d_carray* h, size_t esize)
    {
        fd2d_carray d{nullptr, {0, 0, 1}};
        if (!h) return d;
        d.dims[0] = h->dims[0];
        d.dims[1] = h->dims[1];
        d.dims[2] = h->dims[2];
        const int64_t n = numel(h);
        if (n == 0) return d;
        bufs.emplace_back(new DBuf<char>((size_t)n * esize));
        HIPCHK(hipMemcpy(bufs.back()->p, h->data, (size_t)n * esize, hipMemcpyHostToDevice));
        d.data = bufs.back()->p;
        return d;
    }
    fd2d_marray out(const fd2d_marray* h, size_t esize)
    {
        fd2d_marray d{nullptr, {0, 0, 1}};
        if (!h || !h->data) return d;
        d.dims[0] = h->dims[0];
        d.dims[1] = h->dims[1];
        d.dims[2] = h->dims[2];
        const int64_t n = numel(h);
        if (n == 0) return d;
        bufs.emplace_back(new DBuf<char>((size_t)n * esize));
        d.data = bufs.back()->p;
        return d;
    }
    static void back(fd2d_marray* h, const fd2d_marray& d, size_t esize)
    {
        if (!h || !h->data || !d.data) return;
        HIPCHK(hipMemcpy(h->data, d.data, (size_t)numel(h) * esize, hipMemcpyDeviceToHost));
    }
};

const fd2d_carray* ptr_or_null(const fd2d_carray& a) { return a.data ? &a : nullptr; }

// A solve that ends non-finite still returns its outputs (the reference returns whatever its arrays
// hold): `back` copies them to the host before the failure is reported.
template <class F, class B>
void solve_then_copy(F&& solve, B&& back)
{
    try {
        solve();
    } catch (const Failure& e) {
        if (e.code != FD2D_ENONFINITE) throw;
        const std::string msg = g_err;
        back();
        g_err = msg;
        throw;
    }
    back();
}

template <class F>
int guarded(F&& f)
{
    try {
        g_err.clear();
        f();
        return FD2D_OK;
    } catch (const Failure& e) {
        (void)hipGetLastError();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return FD2D_ENOMEM;
    } catch (...) {
        g_err = "unexpected C++ exception";
        return FD2D_EDEVICE;
    }
}

void require_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        raise(FD2D_EDEVICE, "no HIP device is visible");
    }
}

// Interior block of the parameter map: [ix, iz] = find(absbound == 1, 1, 'first')
// (map_m_to_parameters.m:18-19), 0-based, with the block rows ix..nx-1-ix, columns iz..nz-1-iz.
struct MapFrame {
    int ix, iz, n1, n2;
    std::vector<double> x1, x2;  // coordinates of the block (define_computational_domain.m:51-52)
};

MapFrame map_frame(const fd2d_config* cfg, const Geometry& g)
{
    std::vector<double> ax, az;
    taper_d(cfg, g, ax, az);
    MapFrame f{-1, -1, 0, 0, {}, {}};
    // the first node of absbound == 1 in column-major order: lowest column with az == 1, then the
    // lowest row with ax == 1 (absbound = ax * az, each <= 1, is 1 only where both are)
    for (int j = 0; j < g.nz && f.iz < 0; ++j)
        if (az[(size_t)j] == 1.0) f.iz = j;
    for (int i = 0; i < g.nx && f.ix < 0; ++i)
        if (ax[(size_t)i] == 1.0) f.ix = i;
    if (f.ix < 0 || f.iz < 0) raise(FD2D_EINVAL, "the absorbing frame leaves no node with absbound == 1");
    f.n1 = g.nx - 2 * f.ix;
    f.n2 = g.nz - 2 * f.iz;
    if (f.n1 < 1 || f.n2 < 1) raise(FD2D_EINVAL, "the absorbing frame leaves an empty interior block");
    for (int i = 0; i < f.n1; ++i) f.x1.push_back((f.ix + i) * g.dx);
    for (int j = 0; j < f.n2; ++j) f.x2.push_back((f.iz + j) * g.dz);
    return f;
}

void check_sigma(const double* s, const char* what)
{
    if (!s || !(s[0] > 0) || !(s[1] > 0)) raise(FD2D_EINVAL, "%s must hold two positive widths", what);
}

// Blurs the interior block of each of n_maps nx x nz maps of src into dst (dst may equal src).
void blur_maps(const Geometry& g, const MapFrame& f, const double* src, double* dst, int n_maps,
               const double* sig_src, const double* sig_str, hipStream_t st)
{
    DBuf<double> x1, x2, tmp((size_t)f.n1 * f.n2);
    upload(x1, f.x1, st);
    upload(x2, f.x2, st);
    const size_t off = (size_t)f.ix + (size_t)g.nx * f.iz;
    for (int q = 0; q < n_maps; ++q) {
        const double* sg = q < n_maps - 1 ? sig_src : sig_str;
        launch_gaussblur(src + g.P * q + off, g.nx, dst + g.P * q + off, g.nx, tmp.p, x1.p, x2.p, f.n1, f.n2, sg[0],
                         sg[1], st);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
}

}  // namespace

struct fd2d_engine {
    Engine e;
    explicit fd2d_engine(hipStream_t s) : e(s) {}
};

// A device-resident array owned by the caller through its handle.
struct fd2d_resident {
    void* p = nullptr;
    int64_t dims[3] = {0, 0, 0};
    int32_t elem_bytes = 0;
    size_t bytes = 0;
    ~fd2d_resident()
    {
        if (p) (void)hipFree(p);
    }
};

namespace {
// validated before any device call, so argument errors are reported without a GPU
size_t resident_bytes(const int64_t dims[3], int32_t eb)
{
    if (eb != 4 && eb != 8 && eb != 16) raise(FD2D_EINVAL, "elem_bytes must be 4, 8 or 16, not %d", eb);
    double n = eb;
    for (int q = 0; q < 3; ++q) {
        if (dims[q] < 0) raise(FD2D_EINVAL, "negative dimension");
        n *= (double)dims[q];
    }
    if (n > 4.0e18) raise(FD2D_EINVAL, "array too large");
    return (size_t)dims[0] * (size_t)dims[1] * (size_t)dims[2] * (size_t)eb;
}

std::unique_ptr<fd2d_resident> make_resident(const int64_t dims[3], int32_t eb)
{
    const size_t b = resident_bytes(dims, eb);
    require_device();
    std::unique_ptr<fd2d_resident> r(new fd2d_resident);
    for (int q = 0; q < 3; ++q) r->dims[q] = dims[q];
    r->elem_bytes = eb;
    r->bytes = b;
    if (b && hipMalloc(&r->p, b) != hipSuccess) {
        (void)hipGetLastError();
        r->p = nullptr;
        raise(FD2D_ENOMEM, "device allocation of a %zu-byte resident array failed", b);
    }
    return r;
}
}  // namespace

namespace {
// Scratch of fd2d_make_adjoint_sources_noalloc: a static device array, so that entry point allocates
// nothing. Each device holds its own copy from the library's load until the process ends; host code
// reaches it only through the per-device address hipGetSymbolAddress returns (the array's name is a
// host-side stand-in there, not a device address). The kernels write every element of it they read.
constexpr int64_t kNoAllocScratchDoubles = (int64_t)1 << 21;  // 16 MiB: 3 * nr * nt doubles
static __device__ __align__(256) double noalloc_scratch[kNoAllocScratchDoubles];

std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kMaxDevices] = {};

// The current device's address of noalloc_scratch, looked up once per device (on its first call).
double* noalloc_scratch_ptr()
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) raise(FD2D_EDEVICE, "device %d is beyond the library's %d devices", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
        if (!p) raise(FD2D_EDEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        g_noalloc_ptr[dev] = (double*)p;
    }
    return g_noalloc_ptr[dev];
}

// noalloc: take the kernels' scratch from noalloc_scratch instead of allocating it, and read the
// window limits straight from t_host (no host copy either)
int adjoint_sources_entry(const double* u, const double* u0, const double* du, const double* t, const double* t_host,
                          int64_t nr, int64_t nt, int32_t veldis, int32_t measurement, int32_t deriv_order,
                          double win_left, double win_right, double* misfit, double* adstf, void* stream, bool wait,
                          bool noalloc = false)
{
    return guarded([&] {
        require_device();
        if (nr < 1 || nt < 2 || !u || !u0 || !t || !misfit || !adstf) raise(FD2D_EINVAL, "u, u_0, t, misfit, adstf and nt >= 2 are required");
        if (measurement < 0 || measurement > 3) raise(FD2D_EINVAL, "unknown measurement %d", measurement);
        if (deriv_order != 1 && deriv_order != 2) raise(FD2D_EINVAL, "deriv_order must be 1 or 2");
        if (measurement == FD2D_CC_TIME_SHIFT && deriv_order == 2)
            raise(FD2D_EUNSUPPORTED, "no Hessian vector products for traveltime measurements (make_adjoint_sources.m:121)");
        if (deriv_order == 2 && !du) raise(FD2D_EINVAL, "deriv_order 2 needs du");
        if (noalloc && 3 * nr * nt > kNoAllocScratchDoubles)
            raise(FD2D_EINVAL, "nr x nt = %lld x %lld needs %lld doubles of scratch; the no-allocation entry point "
                  "holds %lld (nr * nt <= %lld)", (long long)nr, (long long)nt, (long long)(3 * nr * nt),
                  (long long)kNoAllocScratchDoubles, (long long)(kNoAllocScratchDoubles / 3));
        hipStream_t st = as_stream(stream);
        StreamScope scope(st);
        // make_adjoint_sources.m:85-99: window limits against the time axis
        std::vector<double> th_copy;
        if (!t_host) th_copy = to_host<double>(t, (size_t)nt);
        const double* th = t_host ? t_host : th_copy.data();
        double left = win_left, right = win_right;
        if (measurement != FD2D_WAVEFORM_DIFFERENCE) {
            if (left < 0) {
                int64_t k = 0;
                while (k < nt && th[(size_t)k] != 0.0) ++k;
                if (k + 1 >= nt) raise(FD2D_EINVAL, "time axis has no t == 0 sample");
                left = th[(size_t)k + 1];
            }
            if (right > th[(size_t)nt - 1]) right = th[(size_t)nt - 1];
            if (left > th[(size_t)nt - 1]) raise(FD2D_EINVAL, "time series too short? (make_adjoint_sources.m:99)");
        }
        DBuf<double> work;
        if (!noalloc) work.alloc((size_t)3 * nr * nt);
        AdjArgs a{u, u0, du, t, nr, nt, th[1] - th[0] >= 0 ? th[1] - th[0] : th[0] - th[1], veldis ? 1 : 0,
                  measurement, deriv_order, left, right, misfit, adstf, noalloc ? noalloc_scratch_ptr() : work.p};
        launch_adjoint_sources(a, st);
        HIPCHK(hipGetLastError());
        if (wait) HIPCHK(hipStreamSynchronize(st));
    });
}

}  // namespace

extern "C" {

const char* fd2d_last_error(void) { return g_err.c_str(); }

int fd2d_set_steps_per_launch(int32_t steps)
{
    if (steps != 0 && steps != 1 && steps != 2) {
        g_err = "steps per launch must be 0 (tuned per shape), 1 or 2";
        return FD2D_EINVAL;
    }
    g_steps_per_launch.store(steps);
    return FD2D_OK;
}

int fd2d_resident_upload(const fd2d_carray* host, int32_t elem_bytes, fd2d_resident** out)
{
    return guarded([&] {
        if (!host || !out) raise(FD2D_EINVAL, "host and out are required");
        const size_t b = resident_bytes(host->dims, elem_bytes);
        if (b && !host->data) raise(FD2D_EINVAL, "host data is NULL");
        std::unique_ptr<fd2d_resident> r = make_resident(host->dims, elem_bytes);
        if (b) HIPCHK(hipMemcpy(r->p, host->data, b, hipMemcpyHostToDevice));
        *out = r.release();
    });
}

int fd2d_resident_alloc(const int64_t dims[3], int32_t elem_bytes, fd2d_resident** out)
{
    return guarded([&] {
        if (!dims || !out) raise(FD2D_EINVAL, "dims and out are required");
        std::unique_ptr<fd2d_resident> r = make_resident(dims, elem_bytes);
        if (r->bytes) HIPCHK(hipMemset(r->p, 0, r->bytes));
        *out = r.release();
    });
}

int fd2d_resident_view(const fd2d_resident* r, fd2d_carray* in, fd2d_marray* out)
{
    return guarded([&] {
        if (!r) raise(FD2D_EINVAL, "resident array is NULL");
        if (in) {
            in->data = r->p;
            for (int q = 0; q < 3; ++q) in->dims[q] = r->dims[q];
        }
        if (out) {
            out->data = r->p;
            for (int q = 0; q < 3; ++q) out->dims[q] = r->dims[q];
        }
    });
}

int fd2d_resident_download(const fd2d_resident* r, fd2d_marray* host, void* stream)
{
    return guarded([&] {
        if (!r || !host) raise(FD2D_EINVAL, "resident array and host are required");
        const size_t b = resident_bytes(host->dims, r->elem_bytes);
        if (b != r->bytes) raise(FD2D_EINVAL, "host array holds %zu bytes, the resident array %zu", b, r->bytes);
        if (!b) return;
        if (!host->data) raise(FD2D_EINVAL, "host data is NULL");
        HIPCHK(hipStreamSynchronize(as_stream(stream)));
        HIPCHK(hipMemcpy(host->data, r->p, b, hipMemcpyDeviceToHost));
    });
}

int fd2d_resident_free(fd2d_resident* r)
{
    return guarded([&] {
        if (!r) return;
        HIPCHK(hipDeviceSynchronize());
        delete r;
    });
}

int fd2d_release_cached_memory(void)
{
    return guarded([&] {
        HIPCHK(hipDeviceSynchronize());
        for (BlockCache& c : g_cache) c.clear();
        std::lock_guard<std::mutex> lk(g_table_mu);
        for (int d = 0; d < 64; ++d) {  // phase tables that no solve holds
            TableCache& c = g_tables[d];
            for (size_t q = 0; q < c.lru.size();) {
                if (c.lru[q].second.use_count() == 1) {
                    c.bytes -= 3 * c.lru[q].second->n * sizeof(unsigned short);
                    c.lru.erase(c.lru.begin() + (long)q);
                } else {
                    ++q;
                }
            }
        }
    });
}

int fd2d_engine_create(void* stream, fd2d_engine** engine)
{
    return guarded([&] {
        if (!engine) raise(FD2D_EINVAL, "engine is NULL");
        require_device();
        std::unique_ptr<fd2d_engine> e(new fd2d_engine(as_stream(stream)));
        *engine = e.release();
    });
}

int fd2d_engine_submit(fd2d_engine* engine, const fd2d_solve* solve, int64_t* ticket)
{
    return guarded([&] {
        if (!engine) raise(FD2D_EINVAL, "engine is NULL");
        const int64_t t = engine->e.submit(solve);
        if (ticket) *ticket = t;
    });
}

int fd2d_engine_advance(fd2d_engine* engine, int32_t rounds, int32_t* n_active)
{
    return guarded([&] {
        if (!engine) raise(FD2D_EINVAL, "engine is NULL");
        if (rounds < 0) raise(FD2D_EINVAL, "rounds must be >= 0");
        engine->e.advance(rounds);
        if (n_active) *n_active = (int32_t)engine->e.pending();
    });
}

int fd2d_engine_done(const fd2d_engine* engine, int64_t ticket, int32_t* done)
{
    return guarded([&] {
        if (!engine || !done) raise(FD2D_EINVAL, "engine and done must not be NULL");
        if (ticket < 1 || ticket >= engine->e.next_ticket) raise(FD2D_EINVAL, "unknown ticket %lld", (long long)ticket);
        *done = const_cast<fd2d_engine*>(engine)->e.is_done(ticket) ? 1 : 0;
    });
}

int fd2d_engine_progress(const fd2d_engine* engine, int64_t ticket, int32_t* state, int64_t* steps_issued,
                         int64_t* steps_total)
{
    return guarded([&] {
        if (!engine || !state) raise(FD2D_EINVAL, "engine and state must not be NULL");
        if (ticket < 1 || ticket >= engine->e.next_ticket) raise(FD2D_EINVAL, "unknown ticket %lld", (long long)ticket);
        int64_t a = 0, b = 0;
        *state = const_cast<fd2d_engine*>(engine)->e.state_of(ticket, &a, &b);
        if (steps_issued) *steps_issued = a;
        if (steps_total) *steps_total = b;
    });
}

int fd2d_engine_cancel(fd2d_engine* engine, int64_t ticket, int32_t* cancelled)
{
    return guarded([&] {
        if (!engine) raise(FD2D_EINVAL, "engine is NULL");
        if (ticket < 1 || ticket >= engine->e.next_ticket) raise(FD2D_EINVAL, "unknown ticket %lld", (long long)ticket);
        StreamScope scope(engine->e.st);
        const bool c = engine->e.cancel(ticket);
        if (cancelled) *cancelled = c ? 1 : 0;
    });
}

int fd2d_engine_destroy(fd2d_engine* engine)
{
    return guarded([&] { delete engine; });
}

int fd2d_run_batch_device(const fd2d_solve* solves, int32_t n, void* stream)
{
    return guarded([&] {
        if (n < 0 || (n > 0 && !solves)) raise(FD2D_EINVAL, "solves is NULL");
        require_device();
        Engine e(as_stream(stream));
        std::vector<int64_t> tk((size_t)n);
        for (int32_t q = 0; q < n; ++q) tk[(size_t)q] = e.submit(&solves[q]);
        while (!e.act.empty()) e.advance(1 << 20);
        e.sync_lanes();
        HIPCHK(hipStreamSynchronize(e.st));
        e.reap();
        std::string bad;
        for (int32_t q = 0; q < n; ++q) {
            int64_t a = 0, b = 0;
            if (e.state_of(tk[(size_t)q], &a, &b) == FD2D_SOLVE_FAILED)
                bad += (bad.empty() ? "" : ", ") + std::to_string(q);
        }
        if (!bad.empty())
            raise(FD2D_ENONFINITE, "solve(s) %s of the batch became non-finite (NaN or Inf): the time step may "
                  "exceed the stability (CFL) limit of the FD order; the other solves are complete", bad.c_str());
    });
}

int fd2d_solve_device_bytes(const fd2d_solve* solve, int64_t* bytes)
{
    return guarded([&] {
        if (!bytes) raise(FD2D_EINVAL, "bytes is NULL");
        *bytes = solve_bytes(solve);
    });
}

int fd2d_timing_enable(int32_t every)
{
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_timing_every = every > 0 ? every : 0;
    g_timing = fd2d_timing{};
    return FD2D_OK;
}

int fd2d_timing_read(fd2d_timing* out)
{
    if (!out) return FD2D_EINVAL;
    std::lock_guard<std::mutex> lk(g_timing_mu);
    *out = g_timing;
    g_timing = fd2d_timing{};
    return FD2D_OK;
}
const char* fd2d_version(void) { return "fd2d_noise_amd 0.1 (gfx950)"; }

int fd2d_run_forward1_green_device(const fd2d_config* cfg, const fd2d_carray* mu, const fd2d_carray* rho,
                                   const fd2d_carray* src, const fd2d_carray* rec, double mode, const fd2d_carray* dmu,
                                   const fd2d_carray* G_in, fd2d_marray* G_fft, fd2d_marray* G_out,
                                   fd2d_marray* seismograms, void* stream)
{
    return guarded([&] {
        require_device();
        green_device(cfg, mu, rho, src, rec, mode, dmu, G_in, G_fft, G_out, seismograms, as_stream(stream));
    });
}

int fd2d_run_forward2_correlation_device(const fd2d_config* cfg, const fd2d_carray* mu, const fd2d_carray* rho,
                                         const fd2d_carray* G_fft, const fd2d_carray* spectrum,
                                         const fd2d_carray* source_dist, const fd2d_carray* rec, double mode,
                                         const fd2d_carray* dmu, const fd2d_carray* C_in, fd2d_marray* seismograms,
                                         fd2d_marray* C_out, void* stream)
{
    return guarded([&] {
        require_device();
        correlation_device(cfg, mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu, C_in, seismograms, C_out,
                           as_stream(stream));
    });
}

int fd2d_run_noise_adjoint_device(const fd2d_config* cfg, const fd2d_carray* mu, const fd2d_carray* rho,
                                  const fd2d_carray* u_fwd, const fd2d_carray* stf, const fd2d_carray* src,
                                  const fd2d_carray* rec, const fd2d_carray* spectrum, const fd2d_carray* source_dist,
                                  const fd2d_carray* G_fft, double mode, const fd2d_carray* dmu,
                                  const fd2d_carray* u_in, fd2d_marray* K, fd2d_marray* stf_fft, fd2d_marray* u_out,
                                  fd2d_marray* seismograms, void* stream)
{
    return guarded([&] {
        require_device();
        adjoint_device(cfg, mu, rho, u_fwd, stf, src, rec, spectrum, source_dist, G_fft, mode, dmu, u_in, K, stf_fft,
                       u_out, seismograms, as_stream(stream));
    });
}

int fd2d_run_forward1_green(const fd2d_config* cfg, const fd2d_carray* mu, const fd2d_carray* rho,
                            const fd2d_carray* src, const fd2d_carray* rec, double mode, const fd2d_carray* dmu,
                            const fd2d_carray* G_in, fd2d_marray* G_fft, fd2d_marray* G_out, fd2d_marray* seismograms)
{
    return guarded([&] {
        require_device();
        Staged s;
        fd2d_carray dmu_d = s.in(dmu, 8), mu_d = s.in(mu, 8), rho_d = s.in(rho, 8), src_d = s.in(src, 8),
                    rec_d = s.in(rec, 8), gin_d = s.in(G_in, 4);
        fd2d_marray gf = s.out(G_fft, 16), go = s.out(G_out, 4), sg = s.out(seismograms, 8);
        solve_then_copy(
            [&] {
                green_device(cfg, &mu_d, &rho_d, ptr_or_null(src_d), ptr_or_null(rec_d), mode, ptr_or_null(dmu_d),
                             ptr_or_null(gin_d), &gf, &go, &sg, nullptr);
            },
            [&] {
                Staged::back(G_fft, gf, 16);
                if (mode != 0) Staged::back(G_out, go, 4);
                Staged::back(seismograms, sg, 8);
            });
    });
}

int fd2d_run_forward2_correlation(const fd2d_config* cfg, const fd2d_carray* mu, const fd2d_carray* rho,
                                  const fd2d_carray* G_fft, const fd2d_carray* spectrum,
                                  const fd2d_carray* source_dist, const fd2d_carray* rec, double mode,
                                  const fd2d_carray* dmu, const fd2d_carray* C_in, fd2d_marray* seismograms,
                                  fd2d_marray* C_out)
{
    return guarded([&] {
        require_device();
        Staged s;
        fd2d_carray mu_d = s.in(mu, 8), rho_d = s.in(rho, 8), g_d = s.in(G_fft, 16), sp_d = s.in(spectrum, 8),
                    sd_d = s.in(source_dist, 8), rec_d = s.in(rec, 8), dmu_d = s.in(dmu, 8), cin_d = s.in(C_in, 4);
        fd2d_marray sg = s.out(seismograms, 8), co = s.out(C_out, 4);
        solve_then_copy(
            [&] {
                correlation_device(cfg, &mu_d, &rho_d, &g_d, &sp_d, &sd_d, ptr_or_null(rec_d), mode,
                                   ptr_or_null(dmu_d), ptr_or_null(cin_d), &sg, &co, nullptr);
            },
            [&] {
                Staged::back(seismograms, sg, 8);
                if (mode != 0) Staged::back(C_out, co, 4);
            });
    });
}

int fd2d_run_noise_adjoint(const fd2d_config* cfg, const fd2d_carray* mu, const fd2d_carray* rho,
                           const fd2d_carray* u_fwd, const fd2d_carray* stf, const fd2d_carray* src,
                           const fd2d_carray* rec, const fd2d_carray* spectrum, const fd2d_carray* source_dist,
                           const fd2d_carray* G_fft, double mode, const fd2d_carray* dmu, const fd2d_carray* u_in,
                           fd2d_marray* K, fd2d_marray* stf_fft, fd2d_marray* u_out, fd2d_marray* seismograms)
{
    return guarded([&] {
        require_device();
        Staged s;
        fd2d_carray mu_d = s.in(mu, 8), rho_d = s.in(rho, 8), uf_d = s.in(u_fwd, 4), stf_d = s.in(stf, 16),
                    src_d = s.in(src, 8), rec_d = s.in(rec, 8), sp_d = s.in(spectrum, 8), sd_d = s.in(source_dist, 8),
                    g_d = s.in(G_fft, 16), dmu_d = s.in(dmu, 8), uin_d = s.in(u_in, 4);
        fd2d_marray kk = s.out(K, 8), sf = s.out(stf_fft, 16), uo = s.out(u_out, 4), sg = s.out(seismograms, 8);
        const int im = (int)mode;
        solve_then_copy(
            [&] {
                adjoint_device(cfg, &mu_d, &rho_d, ptr_or_null(uf_d), ptr_or_null(stf_d), ptr_or_null(src_d),
                               ptr_or_null(rec_d), ptr_or_null(sp_d), ptr_or_null(sd_d), ptr_or_null(g_d), mode,
                               ptr_or_null(dmu_d), ptr_or_null(uin_d), &kk, &sf, &uo, &sg, nullptr);
            },
            [&] {
                Staged::back(K, kk, 8);
                if (im == 1 || im == 12 || im == 13) Staged::back(stf_fft, sf, 16);
                if (im == 11 || im == 13) Staged::back(u_out, uo, 4);
                Staged::back(seismograms, sg, 8);
            });
    });
}

// ---- callers: measurement, inner products, parameter map ------------------------------------
int fd2d_make_adjoint_sources_device(const double* u, const double* u0, const double* du, const double* t, int64_t nr,
                                     int64_t nt, int32_t veldis, int32_t measurement, int32_t deriv_order,
                                     double win_left, double win_right, double* misfit, double* adstf, void* stream)
{
    return adjoint_sources_entry(u, u0, du, t, nullptr, nr, nt, veldis, measurement, deriv_order, win_left, win_right,
                                 misfit, adstf, stream, true);
}

int fd2d_make_adjoint_sources_queued(const double* u, const double* u0, const double* du, const double* t,
                                     const double* t_host, int64_t nr, int64_t nt, int32_t veldis, int32_t measurement,
                                     int32_t deriv_order, double win_left, double win_right, double* misfit,
                                     double* adstf, void* stream)
{
    if (!t_host) {
        g_err = "t_host is required";
        return FD2D_EINVAL;
    }
    return adjoint_sources_entry(u, u0, du, t, t_host, nr, nt, veldis, measurement, deriv_order, win_left, win_right,
                                 misfit, adstf, stream, false);
}

int fd2d_make_adjoint_sources_noalloc(const double* u, const double* u0, const double* du, const double* t,
                                      const double* t_host, int64_t nr, int64_t nt, int32_t veldis,
                                      int32_t measurement, int32_t deriv_order, double win_left, double win_right,
                                      double* misfit, double* adstf, void* stream)
{
    if (!t_host) {
        g_err = "t_host is required";
        return FD2D_EINVAL;
    }
    return adjoint_sources_entry(u, u0, du, t, t_host, nr, nt, veldis, measurement, deriv_order, win_left, win_right,
                                 misfit, adstf, stream, false, true);
}

int fd2d_make_adjoint_sources(const double* u, const double* u0, const double* du, const double* t, int64_t nr,
                              int64_t nt, int32_t veldis, int32_t measurement, int32_t deriv_order, double win_left,
                              double win_right, double* misfit, double* adstf)
{
    int rc = FD2D_OK;
    const int st = guarded([&] {
        require_device();
        const size_t n = (size_t)(nr > 0 ? nr : 0) * (size_t)(nt > 0 ? nt : 0);
        DBuf<double> du_d, u_d(n), u0_d(n), t_d((size_t)(nt > 0 ? nt : 0)), m_d((size_t)(nr > 0 ? nr : 0)), a_d(n);
        if (!u || !u0 || !t || !misfit || !adstf || n == 0) raise(FD2D_EINVAL, "u, u_0, t, misfit, adstf are required");
        HIPCHK(hipMemcpy(u_d.p, u, n * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(u0_d.p, u0, n * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(t_d.p, t, (size_t)nt * 8, hipMemcpyHostToDevice));
        if (du) {
            du_d.alloc(n);
            HIPCHK(hipMemcpy(du_d.p, du, n * 8, hipMemcpyHostToDevice));
        }
        rc = fd2d_make_adjoint_sources_device(u_d.p, u0_d.p, du ? du_d.p : nullptr, t_d.p, nr, nt, veldis, measurement,
                                              deriv_order, win_left, win_right, m_d.p, a_d.p, nullptr);
        if (rc == FD2D_OK) {
            HIPCHK(hipMemcpy(misfit, m_d.p, (size_t)nr * 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(adstf, a_d.p, n * 8, hipMemcpyDeviceToHost));
        }
    });
    return st != FD2D_OK ? st : rc;
}

int fd2d_dot_device(const double* x, const double* y, int64_t n, double* result, void* stream)
{
    return guarded([&] {
        require_device();
        if (!result || n < 0 || (n > 0 && (!x || !y))) raise(FD2D_EINVAL, "dot: null argument");
        hipStream_t st = as_stream(stream);
        StreamScope scope(st);
        const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256));
        DBuf<double> part((size_t)2 * nb + 1);
        launch_dot(x, y, n, part.p, nb, part.p + 2 * nb, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(result, part.p + 2 * nb, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

int fd2d_axpby_device(double alpha, const double* x, double beta, const double* y, double* out, int64_t n, void* stream)
{
    return guarded([&] {
        require_device();
        if (n < 0 || (n > 0 && (!x || !y || !out))) raise(FD2D_EINVAL, "axpby: null argument");
        hipStream_t st = as_stream(stream);
        StreamScope scope(st);
        if (n) launch_axpby(alpha, x, beta, y, out, n, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

int fd2d_axpby_f32_device(float alpha, const float* x, float beta, const float* y, float* out, int64_t n, void* stream)
{
    return guarded([&] {
        require_device();
        if (n < 0 || (n > 0 && (!x || !y || !out))) raise(FD2D_EINVAL, "axpby_f32: null argument");
        hipStream_t st = as_stream(stream);
        StreamScope scope(st);
        if (n) launch_axpby_f32(alpha, x, beta, y, out, n, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

int fd2d_gaussblur2d_device(const double* Y, int32_t n1, int32_t n2, int32_t ldy, const double* x1, const double* x2,
                            double sigma1, double sigma2, double* GY, int32_t ldg, void* stream)
{
    return guarded([&] {
        require_device();
        if (!Y || !GY || !x1 || !x2 || n1 < 1 || n2 < 1 || ldy < n1 || ldg < n1)
            raise(FD2D_EINVAL, "gaussblur2d: bad arguments");
        if (!(sigma1 > 0) || !(sigma2 > 0)) raise(FD2D_EINVAL, "gaussblur2d: sigma must be positive");
        hipStream_t st = as_stream(stream);
        StreamScope scope(st);
        DBuf<double> tmp((size_t)n1 * n2);
        launch_gaussblur(Y, ldy, GY, ldg, tmp.p, x1, x2, n1, n2, sigma1, sigma2, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

int fd2d_general_source_device(const double* spectrum, int32_t n_sample, int32_t n_src, const double* dist,
                               int64_t P, int32_t n_basis_fct, double* dist_basis, void* stream)
{
    return guarded([&] {
        require_device();
        if (!spectrum || !dist || !dist_basis || n_sample < 1 || n_src < 1 || P < 1 || n_basis_fct < 1)
            raise(FD2D_EINVAL, "general_source: bad arguments");
        if (n_basis_fct > n_sample) raise(FD2D_EINVAL, "general_source: n_basis_fct > n_sample leaves empty bands");
        // integration_limits.m:4-12 (1-based limits, stored 0-based)
        std::vector<int> lo((size_t)n_basis_fct), hi((size_t)n_basis_fct);
        const double step = (double)n_sample / (double)n_basis_fct;
        for (int b = 1; b <= n_basis_fct; ++b) {
            lo[(size_t)b - 1] = (int)std::floor((b - 1) * step + 1) - 1;
            hi[(size_t)b - 1] = (int)std::floor(b * step) - 1;
        }
        hipStream_t st = as_stream(stream);
        StreamScope scope(st);
        DBuf<int> lo_d, hi_d;
        upload(lo_d, lo, st);
        upload(hi_d, hi, st);
        launch_general_source(spectrum, n_sample, n_src, dist, P, lo_d.p, hi_d.p, n_basis_fct, dist_basis, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

int fd2d_map_m_to_parameters_device(const fd2d_config* cfg, const double* m, int32_t n_maps,
                                    const double sigma_source[2], const double sigma_structure[2],
                                    double ref_source, double ref_structure, double mu_0, double* m_parameters,
                                    void* stream)
{
    return guarded([&] {
        require_device();
        const Geometry g = geometry(cfg);
        if (!m || !m_parameters || n_maps < 1) raise(FD2D_EINVAL, "m, m_parameters and n_maps >= 1 are required");
        check_sigma(sigma_source, "sigma_source");
        check_sigma(sigma_structure, "sigma_structure");
        const MapFrame f = map_frame(cfg, g);
        hipStream_t st = as_stream(stream);
        StreamScope scope(st);
        const size_t n = g.P * (size_t)n_maps;
        // outside the frame the maps pass through unchanged (map_m_to_parameters.m:16)
        if (m_parameters != m) HIPCHK(hipMemcpyAsync(m_parameters, m, n * 8, hipMemcpyDeviceToDevice, st));
        blur_maps(g, f, m, m_parameters, n_maps, sigma_source, sigma_structure, st);
        // :48-49
        if (n_maps > 1) launch_affine(m_parameters, (int64_t)(g.P * (size_t)(n_maps - 1)), 1.0, ref_source, st);
        launch_affine(m_parameters + g.P * (size_t)(n_maps - 1), (int64_t)g.P, mu_0, ref_structure, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

int fd2d_map_gradparameters_to_gradm_device(const fd2d_config* cfg, const double* grad_parameters, int32_t n_maps,
                                            const double sigma_source[2], const double sigma_structure[2],
                                            double mu_0, double* grad_m, void* stream)
{
    return guarded([&] {
        require_device();
        const Geometry g = geometry(cfg);
        if (!grad_parameters || !grad_m || n_maps < 1)
            raise(FD2D_EINVAL, "grad_parameters, grad_m and n_maps >= 1 are required");
        if (grad_parameters == grad_m) raise(FD2D_EINVAL, "grad_m must not alias grad_parameters");
        check_sigma(sigma_source, "sigma_source");
        check_sigma(sigma_structure, "sigma_structure");
        const MapFrame f = map_frame(cfg, g);
        hipStream_t st = as_stream(stream);
        StreamScope scope(st);
        const int64_t n = (int64_t)(g.P * (size_t)n_maps);
        // temp = 0.0 * grad_parameters (map_gradparameters_to_gradm.m:70), then the blurred frame
        launch_axpby(0.0, grad_parameters, 0.0, grad_parameters, grad_m, n, st);
        blur_maps(g, f, grad_parameters, grad_m, n_maps, sigma_source, sigma_structure, st);
        launch_affine(grad_m + g.P * (size_t)(n_maps - 1), (int64_t)g.P, mu_0, 0.0, st);  // :101
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

int fd2d_regularization_device(double* j, double* g, const double* m, const double* m0, int64_t n, int64_t n_struct,
                               const double* w, int64_t n_w, double alpha, double beta, void* stream)
{
    return guarded([&] {
        require_device();
        if (!j || !m || !m0 || !w || n < 1 || n_struct < 1 || n_struct > n || n_w < 1)
            raise(FD2D_EINVAL, "regularization: bad arguments");
        const int64_t n_src = n - n_struct;
        if (n_src % n_w != 0 || (n_w != 1 && n_w != n_struct))
            raise(FD2D_EINVAL, "weighting must hold 1 or nx*nz values (regularization.m:115)");
        hipStream_t st = as_stream(stream);
        StreamScope scope(st);
        const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256));
        DBuf<double> part((size_t)4 * nb + 2);
        launch_regularization(m, m0, w, n_w, n, n_src, alpha, beta, g, part.p, nb, part.p + 4 * nb, st);
        HIPCHK(hipGetLastError());
        double sums[2];
        HIPCHK(hipMemcpyAsync(sums, part.p + 4 * nb, sizeof sums, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        // regularization.m:114-120, in the reference's order
        *j = *j + alpha / 2 * sums[0];
        *j = *j + beta / 2 * sums[1];
    });
}

int fd2d_query_output_dims(const fd2d_config* cfg, int32_t solver, double mode, int64_t n_rec, int64_t dims[4][3])
{
    int count = 0;
    const int st = guarded([&] {
        const Geometry g = geometry(cfg);
        const int64_t N = 2 * (int64_t)g.nt - 1;
        for (int q = 0; q < 4; ++q) dims[q][0] = dims[q][1] = dims[q][2] = 0;
        auto set = [&](int q, int64_t a, int64_t b, int64_t c) {
            dims[q][0] = a;
            dims[q][1] = b;
            dims[q][2] = c;
        };
        if (solver == FD2D_GREEN) {
            set(0, g.nx, g.nz, g.K);
            if (mode != 0) set(1, g.nx, g.nz, green_n_fw(g));
            set(2, n_rec, g.nt, 1);
            count = 3;
        } else if (solver == FD2D_CORRELATION) {
            set(0, n_rec, N, 1);
            if (mode != 0) set(1, g.nx, g.nz, N / g.f);
            count = 2;
        } else if (solver == FD2D_ADJOINT) {
            const int im = (int)mode;
            const int nb = cfg->n_basis_fct;
            set(0, g.nx, g.nz, (nb == 0 ? 1 : nb) + 1);
            if (im == 1 || im == 12 || im == 13) set(1, g.nx, g.nz, g.K);
            if (im == 11 || im == 13) set(2, g.nx, g.nz, N / g.f);
            set(3, n_rec, N, 1);
            count = 4;
        } else {
            raise(FD2D_EINVAL, "unknown solver %d", solver);
        }
    });
    return st == FD2D_OK ? count : -st;
}

}  // extern "C"
