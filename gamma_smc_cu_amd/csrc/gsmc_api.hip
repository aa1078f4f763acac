// Runtime and C ABI of the gfx950 Gamma-SMC decoder (include/gamma_smc_amd.h).
//
// A context owns, on one device: the bitpacked haplotype rows, the per-site
// gap descriptors, a shared reference to the transition tables, and
// reusable scratch (checkpoints, launch metadata, output slots, pinned
// staging). Tables are shared between contexts with the same
// (device, flow field, steps, scaled rates, precision) and freed with the
// last context that uses them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gamma_smc_amd.h"
#include "fb_kernels.h"
#include "tables.h"

namespace gsmc {
namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_OK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw Error(e_ == hipErrorOutOfMemory ? GSMC_ENOMEM : GSMC_EHIP,                  \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                   \
    } while (0)

template <class F> int guarded(F&& f)
{
    try {
        f();
        g_err.clear();
        return GSMC_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return GSMC_ENOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return GSMC_EINVAL;
    }
}

void require(bool ok, const std::string& msg)
{
    if (!ok) throw Error(GSMC_EINVAL, msg);
}

// ------------------------------------------------------------- device buffers
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int device = -1;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) {
            (void)hipSetDevice(device);
            (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t n, int dev)
    {
        if (n <= bytes && p) return;
        // Growing a buffer that queued work may still read: let that work finish.
        if (p) HIP_OK(hipDeviceSynchronize());
        release();
        device = dev;
        HIP_OK(hipMalloc(&p, std::max<size_t>(n, 256)));
        bytes = std::max<size_t>(n, 256);
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf()
    {
        if (p) (void)hipHostFree(p);
    }
    void ensure(size_t n)
    {
        if (n <= bytes && p) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        HIP_OK(hipHostMalloc(&p, n, hipHostMallocDefault));
        bytes = n;
    }
};

// ------------------------------------------------------------- table cache
struct DeviceTables {
    int device = -1;
    int n_steps = 0;
    char* base = nullptr;  // fwd layers, bwd layers, stretch slab
    size_t bytes = 0;
    DevTables view{};
    ~DeviceTables()
    {
        if (base) {
            (void)hipSetDevice(device);
            (void)hipFree(base);
        }
    }
};

std::mutex g_tab_mu;
std::map<std::string, std::weak_ptr<DeviceTables>> g_tabs;

std::string table_key(int device, const FlowField& ff, int N, float rho, float mu, double rho64, double mu64,
                      int precision)
{
    std::string k;
    auto put = [&k](const void* p, size_t n) { k.append(static_cast<const char*>(p), n); };
    put(&device, sizeof device);
    put(&N, sizeof N);
    put(&precision, sizeof precision);
    put(&rho, sizeof rho);
    put(&mu, sizeof mu);
    if (precision == kTablesF64) {
        put(&rho64, sizeof rho64);
        put(&mu64, sizeof mu64);
    }
    put(ff.bounds, sizeof ff.bounds);
    put(ff.u, sizeof ff.u);
    put(ff.v, sizeof ff.v);
    return k;
}

std::shared_ptr<DeviceTables> get_tables(int device, const char* path, int N, double Ne, double mu, double rho,
                                         int precision)
{
    require(path != nullptr, "flow_field_path is required");
    require(precision == kTablesF32 || precision == kTablesF64, "table_precision must be 0 (fp32) or 1 (fp64)");
    require(N >= 1 && N <= (int)kDescLayerMask + 1, "cache_steps must be in [1, 8192]");
    FlowField ff;
    std::string err;
    if (!load_flow_field(path, ff, err)) throw Error(GSMC_EIO, err);
    // FP32 scaled rates exactly as the reference forms them: float Ne, mu, rho,
    // then 4.0f * Ne * rate (bindings.cpp:2575, :2044-2045).
    const float ne_f = (float)Ne, mu_f = (float)mu, rho_f = (float)rho;
    const float srho = 4.0f * ne_f * rho_f, smu = 4.0f * ne_f * mu_f;
    const double srho64 = 4.0 * Ne * rho, smu64 = 4.0 * Ne * mu;
    const std::string key = table_key(device, ff, N, srho, smu, srho64, smu64, precision);
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_tabs.find(key);
    if (it != g_tabs.end()) {
        if (auto sp = it->second.lock()) return sp;
    }
    HostTables host;
    build_tables(ff, N, srho, smu, srho64, smu64, precision, host);
    auto dt = std::make_shared<DeviceTables>();
    dt->device = device;
    dt->n_steps = N;
    const size_t layers = (size_t)N * kLayerBytes;
    dt->bytes = 2 * layers + kSlabBytes;
    HIP_OK(hipMalloc(&dt->base, dt->bytes));
    HIP_OK(hipMemcpy(dt->base, host.fwd.data(), layers, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dt->base + layers, host.bwd.data(), layers, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dt->base + 2 * layers, host.stretch_last.data(), kSlabBytes, hipMemcpyHostToDevice));
    dt->view.fwd = dt->base;
    dt->view.bwd = dt->base + layers;
    dt->view.stretch = dt->base + 2 * layers;
    g_tabs[key] = dt;
    for (auto i = g_tabs.begin(); i != g_tabs.end();) i = i->second.expired() ? g_tabs.erase(i) : std::next(i);
    return dt;
}

// ------------------------------------------------------------- site descriptors
// __double2int_rn: nearest, ties to even, saturating.
int round_steps(double delta)
{
    if (delta != delta) return 1;
    const double r = std::nearbyint(delta);
    int v;
    if (r >= 2147483647.0) v = 2147483647;
    else if (r <= -2147483648.0) v = INT32_MIN;
    else v = (int)r;
    return v > 1 ? v : 1;  // rounded_segment_steps (gamma_smc_flow.cu:160-163)
}

uint32_t descriptor(int steps, int N)
{
    const uint32_t q = (uint32_t)((steps - 1) / N);
    const uint32_t layer = (uint32_t)(steps - (int)q * N - 1);
    if (q >= (1u << (32 - kDescLayerBits)))
        throw Error(GSMC_EINVAL, "gap between consecutive positions too large for cache_steps");
    return layer | (q << kDescLayerBits);
}

}  // namespace
}  // namespace gsmc

using namespace gsmc;

struct gsmc_context {
    std::mutex mu;
    int device = 0;
    int64_t n = 0, S = 0, W = 0;
    float Ne = 0.0f;
    int N = 1000;
    int K = 32;
    int64_t min_waves = 8192;
    std::vector<double> pos;
    DevBuf packed, desc;
    std::shared_ptr<DeviceTables> tabs;
    // launch scratch of the host-output path
    DevBuf ck, meta, pairs;
    // launch scratch of the device-resident path
    DevBuf dev_ck, dev_meta;
    std::vector<char> dev_meta_host;
    DevBuf out_slot[2];
    PinnedBuf stage[2];
    hipStream_t compute = nullptr, copy = nullptr;
    hipEvent_t ev_kernel[2] = {nullptr, nullptr};  // slot written (compute -> copy)
    hipEvent_t ev_free[2] = {nullptr, nullptr};    // slot drained (copy -> compute)
    hipEvent_t ev_stage[2] = {nullptr, nullptr};   // staging buffer filled (copy -> host)

    ~gsmc_context()
    {
        (void)hipSetDevice(device);
        for (auto& e : ev_kernel)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : ev_free)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : ev_stage)
            if (e) (void)hipEventDestroy(e);
        if (compute) (void)hipStreamDestroy(compute);
        if (copy) (void)hipStreamDestroy(copy);
    }
};

namespace gsmc {
namespace {

void init_streams(gsmc_context* ctx)
{
    if (ctx->compute) return;
    HIP_OK(hipStreamCreateWithFlags(&ctx->compute, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&ctx->copy, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
        HIP_OK(hipEventCreateWithFlags(&ctx->ev_kernel[i], hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&ctx->ev_free[i], hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&ctx->ev_stage[i], hipEventDisableTiming));
    }
}

void finish_create(gsmc_context* ctx, const double* positions, double Ne, double mu, double rho, const char* path,
                   int cache_steps, int precision)
{
    ctx->N = cache_steps <= 0 ? 1000 : cache_steps;
    ctx->Ne = (float)Ne;
    ctx->pos.assign(positions, positions + ctx->S);
    ctx->tabs = get_tables(ctx->device, path, ctx->N, Ne, mu, rho, precision);
    std::vector<uint32_t> d((size_t)std::max<int64_t>(ctx->S, 1));
    for (int64_t s = 0; s < ctx->S; ++s) {
        const double delta = s == 0 ? positions[0] + 1.0 : positions[s] - positions[s - 1];
        d[s] = descriptor(round_steps(delta), ctx->N);
    }
    ctx->desc.ensure(d.size() * 4, ctx->device);
    HIP_OK(hipMemcpy(ctx->desc.p, d.data(), d.size() * 4, hipMemcpyHostToDevice));
}

struct Block {
    int32_t cs, ce, ps, pe;
};

std::vector<Block> make_blocks(int64_t S, int64_t core, int64_t flank)
{
    std::vector<Block> out;
    for (int64_t cs = 0; cs < S; cs += core) {
        const int64_t ce = std::min(S, cs + core);
        out.push_back({(int32_t)cs, (int32_t)ce, (int32_t)std::max<int64_t>(0, cs - flank),
                       (int32_t)std::min(S, ce + flank)});
    }
    return out;
}

struct Outputs {
    float* mean;
    float* lower;
    float* upper;
    float* alpha;
    float* beta;
};

// Launch metadata of one decode: block windows (int4), first checkpoint
// window per block (int64), forward descriptor of each block's first padded
// site (uint32). `windows` is the launch's checkpoint row count.
struct Meta {
    std::vector<char> bytes;
    int64_t nb = 0;
    int64_t windows = 0;
};

Meta make_meta(const gsmc_context* ctx, const Block* blocks, int64_t nb)
{
    require(nb <= 65535, "too many block windows for one launch");
    Meta m;
    m.nb = nb;
    m.bytes.resize((size_t)nb * 28);
    std::vector<int64_t> base((size_t)nb);
    std::vector<uint32_t> sdesc((size_t)nb);
    for (int64_t b = 0; b < nb; ++b) {
        base[b] = m.windows;
        m.windows += (blocks[b].ce - blocks[b].cs + ctx->K - 1) / ctx->K;
        const double pps = ctx->pos[blocks[b].ps];
        sdesc[b] = descriptor(round_steps(pps - (pps - 1.0)), ctx->N);
    }
    std::memcpy(m.bytes.data(), blocks, (size_t)nb * 16);
    std::memcpy(m.bytes.data() + (size_t)nb * 16, base.data(), (size_t)nb * 8);
    std::memcpy(m.bytes.data() + (size_t)nb * 24, sdesc.data(), (size_t)nb * 4);
    return m;
}

// Queue one launch over nb blocks whose metadata sits at d_meta, for pairs
// d_pi/d_pj (device) into outputs whose row r holds site row0 + r with row
// stride ld. The checkpoint buffer must hold windows x ckld float2.
void launch_fb(gsmc_context* ctx, const char* d_meta, int64_t nb, float2* ck, const int32_t* d_pi,
               const int32_t* d_pj, int64_t P, int flags, const Outputs& out, int64_t row0, int64_t ld,
               hipStream_t stream)
{
    if (P == 0 || nb == 0) return;
    const int64_t tiles = (P + 63) / 64;
    FbArgs a;
    a.packed = ctx->packed.as<uint64_t>();
    a.W = ctx->W;
    a.desc = ctx->desc.as<uint32_t>();
    a.pair_i = d_pi;
    a.pair_j = d_pj;
    a.P = P;
    a.blocks = reinterpret_cast<const int4*>(d_meta);
    a.ck_base = reinterpret_cast<const int64_t*>(d_meta + (size_t)nb * 16);
    a.start_desc = reinterpret_cast<const uint32_t*>(d_meta + (size_t)nb * 24);
    a.ck = ck;
    a.ckld = tiles * 64;
    a.t = ctx->tabs->view;
    a.unscale = 2.0f * ctx->Ne;
    a.row0 = row0;
    a.ld = ld;
    a.mean = out.mean;
    a.lower = out.lower;
    a.upper = out.upper;
    a.alpha = out.alpha;
    a.beta = out.beta;
    const dim3 grid((unsigned)tiles, (unsigned)nb);
    const bool ci = flags & GSMC_CI, post = flags & GSMC_POSTERIOR;
#define GSMC_LAUNCH(KK, C, Q) hipLaunchKernelGGL((fb_block_kernel<KK, C, Q>), grid, dim3(64), 0, stream, a)
#define GSMC_BY_FLAGS(KK)                        \
    if (ci && post) GSMC_LAUNCH(KK, true, true);   \
    else if (ci) GSMC_LAUNCH(KK, true, false);     \
    else if (post) GSMC_LAUNCH(KK, false, true);   \
    else GSMC_LAUNCH(KK, false, false);
    if (ctx->K == 16) {
        GSMC_BY_FLAGS(16)
    } else {
        GSMC_BY_FLAGS(32)
    }
#undef GSMC_BY_FLAGS
#undef GSMC_LAUNCH
    HIP_OK(hipGetLastError());
}

int n_outputs(int flags) { return 1 + ((flags & GSMC_CI) ? 2 : 0) + ((flags & GSMC_POSTERIOR) ? 2 : 0); }

void check_pairs(const gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P)
{
    for (int64_t p = 0; p < P; ++p)
        require(pi[p] >= 0 && pi[p] < ctx->n && pj[p] >= 0 && pj[p] < ctx->n,
                "pair index out of range: (" + std::to_string(pi[p]) + ", " + std::to_string(pj[p]) + ")");
}

// Parallel host copy of staged rows into the caller's [S][P] arrays.
void scatter_rows(float* dst, int64_t ld_dst, int64_t col0, const float* src, int64_t rows, int64_t cols)
{
    const size_t bytes = (size_t)rows * cols * 4;
    const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes >> 26));
    auto work = [&](int t) {
        const int64_t r0 = rows * t / nt, r1 = rows * (t + 1) / nt;
        if (cols == ld_dst && col0 == 0) {
            std::memcpy(dst + r0 * ld_dst, src + r0 * cols, (size_t)(r1 - r0) * cols * 4);
        } else {
            for (int64_t r = r0; r < r1; ++r) std::memcpy(dst + r * ld_dst + col0, src + r * cols, (size_t)cols * 4);
        }
    };
    if (nt == 1) {
        work(0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
}

// Host-output driver shared by run_fb and run_fb_blockwise: pair chunks x
// groups of consecutive blocks, each group decoded by one launch into a
// device slot, then drained to host through pinned staging while the next
// group runs.
void run_to_host(gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P, const std::vector<Block>& blocks,
                 int64_t pair_batch, int slots, int flags, float* const* host_out)
{
    init_streams(ctx);
    const int nout = n_outputs(flags);
    const int64_t chunk = (pair_batch > 0) ? std::min<int64_t>(pair_batch, P) : P;
    // Blocks per launch: enough pair-tile x block waves to fill the chip,
    // bounded by a device budget for the output slots.
    size_t free_b = 0, total_b = 0;
    HIP_OK(hipMemGetInfo(&free_b, &total_b));
    const int64_t tiles = (chunk + 63) / 64;
    int64_t per_group = std::max<int64_t>(1, (ctx->min_waves + tiles - 1) / tiles);
    int64_t max_rows = 0;
    for (const Block& b : blocks) max_rows = std::max<int64_t>(max_rows, b.ce - b.cs);
    const size_t slot_budget = (size_t)(free_b * 0.30);
    while (per_group > 1 && (size_t)per_group * max_rows * chunk * 4 * nout > slot_budget) per_group /= 2;
    slots = std::max(1, std::min(slots, 2));

    ctx->pairs.ensure((size_t)P * 8, ctx->device);
    int32_t* d_pi = ctx->pairs.as<int32_t>();
    int32_t* d_pj = d_pi + P;
    HIP_OK(hipStreamSynchronize(ctx->compute));
    HIP_OK(hipMemcpy(d_pi, pi, (size_t)P * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_pj, pj, (size_t)P * 4, hipMemcpyHostToDevice));

    struct Task {
        int64_t p0, np, b0, nb, row0, rows;
    };
    std::vector<Task> tasks;
    for (int64_t p0 = 0; p0 < P; p0 += chunk)
        for (int64_t b0 = 0; b0 < (int64_t)blocks.size(); b0 += per_group) {
            const int64_t nb = std::min<int64_t>(per_group, (int64_t)blocks.size() - b0);
            const int64_t r0 = blocks[b0].cs, r1 = blocks[b0 + nb - 1].ce;
            tasks.push_back({p0, std::min(chunk, P - p0), b0, nb, r0, r1 - r0});
        }

    // Metadata of every launch, uploaded once; the largest checkpoint and
    // slot sizes reserved up front, so nothing is reallocated or overwritten
    // while a queued launch can still read it.
    std::vector<size_t> meta_off(tasks.size());
    {
        std::vector<char> all;
        size_t ck_max = 0;
        for (size_t i = 0; i < tasks.size(); ++i) {
            const Task& t = tasks[i];
            Meta m = make_meta(ctx, blocks.data() + t.b0, t.nb);
            meta_off[i] = all.size();
            all.insert(all.end(), m.bytes.begin(), m.bytes.end());
            all.resize((all.size() + 255) & ~(size_t)255);
            ck_max = std::max(ck_max, (size_t)m.windows * (size_t)((t.np + 63) / 64 * 64) * sizeof(float2));
        }
        ctx->ck.ensure(ck_max, ctx->device);
        ctx->meta.ensure(all.size(), ctx->device);
        HIP_OK(hipStreamSynchronize(ctx->compute));
        HIP_OK(hipMemcpy(ctx->meta.p, all.data(), all.size(), hipMemcpyHostToDevice));
        for (int s = 0; s < slots; ++s) {
            size_t slot_max = 0;
            for (size_t i = s; i < tasks.size(); i += slots)
                slot_max = std::max(slot_max, (size_t)tasks[i].rows * tasks[i].np * 4 * nout);
            ctx->out_slot[s].ensure(slot_max, ctx->device);
        }
    }
    bool slot_used[2] = {false, false};
    const size_t stage_bytes = (size_t)256 << 20;
    for (int s = 0; s < 2; ++s) ctx->stage[s].ensure(stage_bytes);
    int stage_next = 0;
    bool stage_busy[2] = {false, false};
    // pending host copies per staging buffer
    struct Pending {
        int out;
        int64_t row, rows, col0, cols;
    } pend[2];

    auto drain_stage = [&](int s) {
        if (!stage_busy[s]) return;
        HIP_OK(hipEventSynchronize(ctx->ev_stage[s]));
        const Pending& q = pend[s];
        scatter_rows(host_out[q.out] + (size_t)q.row * P, P, q.col0, static_cast<const float*>(ctx->stage[s].p),
                     q.rows, q.cols);
        stage_busy[s] = false;
    };

    // copy one finished slot out, through the staging ring, in row pieces
    auto copy_out = [&](int slot, const Task& t) {
        const size_t plane = (size_t)t.rows * t.np;  // floats per output array in the slot
        const int64_t rows_per_piece = std::max<int64_t>(1, (int64_t)(stage_bytes / 4 / std::max<int64_t>(t.np, 1)));
        HIP_OK(hipStreamWaitEvent(ctx->copy, ctx->ev_kernel[slot], 0));
        for (int o = 0; o < nout; ++o) {
            const float* src = ctx->out_slot[slot].as<float>() + (size_t)o * plane;
            for (int64_t r = 0; r < t.rows; r += rows_per_piece) {
                const int64_t nr = std::min(rows_per_piece, t.rows - r);
                const int s = stage_next;
                stage_next ^= 1;
                drain_stage(s);
                HIP_OK(hipMemcpyAsync(ctx->stage[s].p, src + (size_t)r * t.np, (size_t)nr * t.np * 4,
                                      hipMemcpyDeviceToHost, ctx->copy));
                HIP_OK(hipEventRecord(ctx->ev_stage[s], ctx->copy));
                pend[s] = {o, t.row0 + r, nr, t.p0, t.np};
                stage_busy[s] = true;
            }
        }
        HIP_OK(hipEventRecord(ctx->ev_free[slot], ctx->copy));
    };

    int prev_slot = -1;
    const Task* prev_task = nullptr;
    for (size_t i = 0; i < tasks.size(); ++i) {
        const Task& t = tasks[i];
        const int slot = (int)(i % (size_t)slots);
        const size_t plane = (size_t)t.rows * t.np;
        // the slot's previous contents were fully staged before we reuse it
        if (slots == 1 && prev_slot >= 0) {
            copy_out(prev_slot, *prev_task);
            prev_slot = -1;
            for (int s = 0; s < 2; ++s) drain_stage(s);
        }
        if (slot_used[slot]) HIP_OK(hipStreamWaitEvent(ctx->compute, ctx->ev_free[slot], 0));
        slot_used[slot] = true;
        float* base = ctx->out_slot[slot].as<float>();
        Outputs o{};
        int k = 0;
        o.mean = base + plane * k++;
        if (flags & GSMC_CI) {
            o.lower = base + plane * k++;
            o.upper = base + plane * k++;
        }
        if (flags & GSMC_POSTERIOR) {
            o.alpha = base + plane * k++;
            o.beta = base + plane * k++;
        }
        launch_fb(ctx, ctx->meta.as<char>() + meta_off[i], t.nb, ctx->ck.as<float2>(), d_pi + t.p0, d_pj + t.p0,
                  t.np, flags, o, t.row0, t.np, ctx->compute);
        HIP_OK(hipEventRecord(ctx->ev_kernel[slot], ctx->compute));
        if (slots > 1 && prev_slot >= 0) copy_out(prev_slot, *prev_task);
        prev_slot = slot;
        prev_task = &t;
    }
    if (prev_slot >= 0) copy_out(prev_slot, *prev_task);
    drain_stage(stage_next);
    drain_stage(stage_next ^ 1);
    HIP_OK(hipStreamSynchronize(ctx->compute));
    HIP_OK(hipStreamSynchronize(ctx->copy));
}

}  // namespace
}  // namespace gsmc

// ================================================================= C ABI
extern "C" {

const char* gsmc_last_error(void) { return g_err.c_str(); }

int gsmc_build_tables(const char* path, int n_steps, double Ne, double mu, double rho, int precision, float* fwd,
                      float* bwd, float* stretch)
{
    return guarded([&] {
        require(path != nullptr, "flow_field_path is required");
        require(precision == kTablesF32 || precision == kTablesF64, "table_precision must be 0 (fp32) or 1 (fp64)");
        require(n_steps >= 1 && n_steps <= (int)kDescLayerMask + 1, "cache_steps must be in [1, 8192]");
        FlowField ff;
        std::string err;
        if (!load_flow_field(path, ff, err)) throw Error(GSMC_EIO, err);
        const float ne_f = (float)Ne;
        HostTables h;
        build_tables(ff, n_steps, 4.0f * ne_f * (float)rho, 4.0f * ne_f * (float)mu, 4.0 * Ne * rho, 4.0 * Ne * mu,
                     precision, h);
        std::memcpy(fwd, h.fwd.data(), h.fwd.size() * 4);
        std::memcpy(bwd, h.bwd.data(), h.bwd.size() * 4);
        std::memcpy(stretch, h.stretch_last.data(), h.stretch_last.size() * 4);
    });
}

int gsmc_device_count(int* count)
{
    return guarded([&] { HIP_OK(hipGetDeviceCount(count)); });
}

int gsmc_set_device(int device)
{
    return guarded([&] { HIP_OK(hipSetDevice(device)); });
}

int gsmc_mem_info(uint64_t* free_bytes, uint64_t* total_bytes)
{
    return guarded([&] {
        size_t f = 0, t = 0;
        HIP_OK(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
    });
}

int gsmc_bitpack(const uint8_t* G, int64_t n, int64_t S, uint64_t* packed)
{
    return guarded([&] {
        require(n >= 0 && S >= 0, "negative shape");
        const int64_t W = (S + 63) / 64;
        if (n == 0 || W == 0) return;
        DevBuf g, pk;
        int dev = 0;
        HIP_OK(hipGetDevice(&dev));
        g.ensure((size_t)n * S, dev);
        pk.ensure((size_t)n * W * 8, dev);
        HIP_OK(hipMemcpy(g.p, G, (size_t)n * S, hipMemcpyHostToDevice));
        const int64_t waves = n * W;
        const unsigned blocks = (unsigned)std::min<int64_t>((waves + 3) / 4, 1 << 20);
        hipLaunchKernelGGL(bitpack_kernel, dim3(blocks), dim3(256), 0, 0, g.as<uint8_t>(), n, S, W, pk.as<uint64_t>());
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(packed, pk.p, (size_t)n * W * 8, hipMemcpyDeviceToHost));
    });
}

int gsmc_unpack(const uint64_t* packed, int64_t n, int64_t S, uint8_t* G)
{
    return guarded([&] {
        require(n >= 0 && S >= 0, "negative shape");
        const int64_t W = (S + 63) / 64;
        if (n == 0 || S == 0) return;
        DevBuf g, pk;
        int dev = 0;
        HIP_OK(hipGetDevice(&dev));
        g.ensure((size_t)n * S, dev);
        pk.ensure((size_t)n * W * 8, dev);
        HIP_OK(hipMemcpy(pk.p, packed, (size_t)n * W * 8, hipMemcpyHostToDevice));
        const unsigned blocks = (unsigned)std::min<int64_t>((n * S + 255) / 256, 1 << 20);
        hipLaunchKernelGGL(unpack_kernel, dim3(blocks), dim3(256), 0, 0, pk.as<uint64_t>(), n, S, W, g.as<uint8_t>());
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(G, g.p, (size_t)n * S, hipMemcpyDeviceToHost));
    });
}

static int create_common(const uint8_t* G, const uint64_t* packed, int64_t n, int64_t S, const double* positions,
                         double Ne, double mu, double rho, const char* path, int cache_steps, int precision,
                         gsmc_context** out)
{
    gsmc_context* ctx = nullptr;
    int rc = guarded([&] {
        require(out != nullptr, "out is NULL");
        require(n >= 0 && S >= 0, "negative shape");
        require(S == 0 || positions != nullptr, "positions is NULL");
        require(S <= INT32_MAX - 1, "n_sites too large");
        ctx = new gsmc_context();
        HIP_OK(hipGetDevice(&ctx->device));
        ctx->n = n;
        ctx->S = S;
        ctx->W = (S + 63) / 64;
        const size_t pk_bytes = (size_t)n * ctx->W * 8;
        ctx->packed.ensure(pk_bytes, ctx->device);
        if (pk_bytes) {
            if (packed) {
                HIP_OK(hipMemcpy(ctx->packed.p, packed, pk_bytes, hipMemcpyHostToDevice));
            } else {
                DevBuf g;
                g.ensure((size_t)n * S, ctx->device);
                HIP_OK(hipMemcpy(g.p, G, (size_t)n * S, hipMemcpyHostToDevice));
                const int64_t waves = n * ctx->W;
                const unsigned blocks = (unsigned)std::min<int64_t>((waves + 3) / 4, 1 << 20);
                hipLaunchKernelGGL(bitpack_kernel, dim3(blocks), dim3(256), 0, 0, g.as<uint8_t>(), n, S, ctx->W,
                                   ctx->packed.as<uint64_t>());
                HIP_OK(hipGetLastError());
                HIP_OK(hipDeviceSynchronize());
            }
        }
        finish_create(ctx, positions, Ne, mu, rho, path, cache_steps, precision);
    });
    if (rc != GSMC_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return GSMC_OK;
}

int gsmc_context_create(const uint8_t* G, int64_t n, int64_t S, const double* positions, double Ne, double mu,
                        double rho, const char* path, int cache_steps, int precision, gsmc_context** out)
{
    if (G == nullptr && n * S > 0) {
        g_err = "G is NULL";
        return GSMC_EINVAL;
    }
    return create_common(G, nullptr, n, S, positions, Ne, mu, rho, path, cache_steps, precision, out);
}

int gsmc_context_create_packed(const uint64_t* packed, int64_t n, int64_t S, const double* positions, double Ne,
                               double mu, double rho, const char* path, int cache_steps, int precision,
                               gsmc_context** out)
{
    if (packed == nullptr && n * S > 0) {
        g_err = "packed is NULL";
        return GSMC_EINVAL;
    }
    return create_common(nullptr, packed, n, S, positions, Ne, mu, rho, path, cache_steps, precision, out);
}

void gsmc_context_destroy(gsmc_context* ctx) { delete ctx; }

int gsmc_context_info(const gsmc_context* ctx, int64_t* n, int64_t* S, int* device, float* Ne)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        if (n) *n = ctx->n;
        if (S) *S = ctx->S;
        if (device) *device = ctx->device;
        if (Ne) *Ne = ctx->Ne;
    });
}

int gsmc_context_set_tuning(gsmc_context* ctx, int k, int64_t min_waves)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        require(k == 16 || k == 32, "checkpoint_k must be 16 or 32");
        require(min_waves >= 1, "min_waves_per_launch must be positive");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->K = k;
        ctx->min_waves = min_waves;
    });
}

int gsmc_block_windows(int64_t S, int64_t core, int64_t flank, int32_t* blocks, int64_t cap, int64_t* n_blocks)
{
    return guarded([&] {
        require(core > 0, "core_block_sites must be positive.");
        require(flank >= 0, "flank_sites must be non-negative.");
        auto b = make_blocks(S, core, flank);
        if (n_blocks) *n_blocks = (int64_t)b.size();
        for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)b.size()); ++i) {
            blocks[4 * i + 0] = b[i].cs;
            blocks[4 * i + 1] = b[i].ce;
            blocks[4 * i + 2] = b[i].ps;
            blocks[4 * i + 3] = b[i].pe;
        }
    });
}

int gsmc_run_fb(gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P, int flags, float* mean,
                float* lower, float* upper, float* alpha, float* beta)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        require(P >= 0, "negative n_pairs");
        require(mean != nullptr || P * ctx->S == 0, "mean is NULL");
        require(!(flags & GSMC_CI) || (lower && upper) || P * ctx->S == 0, "lower/upper required with GSMC_CI");
        require(!(flags & GSMC_POSTERIOR) || (alpha && beta) || P * ctx->S == 0,
                "alpha/beta required with GSMC_POSTERIOR");
        if (P == 0 || ctx->S == 0) return;
        check_pairs(ctx, pi, pj, P);
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        std::vector<Block> blocks{{0, (int32_t)ctx->S, 0, (int32_t)ctx->S}};
        float* outs[5];
        int k = 0;
        outs[k++] = mean;
        if (flags & GSMC_CI) outs[k++] = lower, outs[k++] = upper;
        if (flags & GSMC_POSTERIOR) outs[k++] = alpha, outs[k++] = beta;
        run_to_host(ctx, pi, pj, P, blocks, -1, 2, flags, outs);
    });
}

int gsmc_run_fb_blockwise(gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P, int64_t core,
                          int64_t flank, int64_t pair_batch, int max_streams, int flags, float* mean, float* lower,
                          float* upper, float* alpha, float* beta)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        require(core > 0, "core_block_sites must be positive.");
        require(flank >= 0, "flank_sites must be non-negative.");
        require(pair_batch > 0 || pair_batch == -1, "pair_batch_size must be positive or -1 for auto.");
        require(max_streams > 0, "max_streams must be positive.");
        require(!((flags & GSMC_POSTERIOR) && max_streams > 1),
                "return_posterior=True is currently only supported with max_streams=1. "
                "Run blockwise posterior decoding in single-stream mode.");
        require(P >= 0, "negative n_pairs");
        if (P == 0 || ctx->S == 0) return;
        require(mean != nullptr, "mean is NULL");
        require(!(flags & GSMC_CI) || (lower && upper), "lower/upper required with GSMC_CI");
        require(!(flags & GSMC_POSTERIOR) || (alpha && beta), "alpha/beta required with GSMC_POSTERIOR");
        check_pairs(ctx, pi, pj, P);
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        auto blocks = make_blocks(ctx->S, core, flank);
        float* outs[5];
        int k = 0;
        outs[k++] = mean;
        if (flags & GSMC_CI) outs[k++] = lower, outs[k++] = upper;
        if (flags & GSMC_POSTERIOR) outs[k++] = alpha, outs[k++] = beta;
        run_to_host(ctx, pi, pj, P, blocks, pair_batch, std::max(2, max_streams), flags, outs);
    });
}

int gsmc_run_fb_device(gsmc_context* ctx, const int32_t* d_pi, const int32_t* d_pj, int64_t P, int64_t core,
                       int64_t flank, int flags, float* d_mean, float* d_lower, float* d_upper, float* d_alpha,
                       float* d_beta, void* stream)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        require(core > 0, "core_block_sites must be positive.");
        require(flank >= 0, "flank_sites must be non-negative.");
        require(P >= 0, "negative n_pairs");
        if (P == 0 || ctx->S == 0) return;
        require(d_mean != nullptr, "mean is NULL");
        require(!(flags & GSMC_CI) || (d_lower && d_upper), "lower/upper required with GSMC_CI");
        require(!(flags & GSMC_POSTERIOR) || (d_alpha && d_beta), "alpha/beta required with GSMC_POSTERIOR");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        auto blocks = make_blocks(ctx->S, core, flank);
        Meta m = make_meta(ctx, blocks.data(), (int64_t)blocks.size());
        const size_t ck_bytes = (size_t)m.windows * (size_t)((P + 63) / 64 * 64) * sizeof(float2);
        // Metadata and checkpoints are reused while the launch shape repeats;
        // a new shape waits for queued work before replacing them.
        if (m.bytes != ctx->dev_meta_host || ck_bytes > ctx->dev_ck.bytes) {
            HIP_OK(hipDeviceSynchronize());
            ctx->dev_meta.ensure(m.bytes.size(), ctx->device);
            ctx->dev_ck.ensure(ck_bytes, ctx->device);
            HIP_OK(hipMemcpy(ctx->dev_meta.p, m.bytes.data(), m.bytes.size(), hipMemcpyHostToDevice));
            ctx->dev_meta_host = m.bytes;
        }
        Outputs o{d_mean, d_lower, d_upper, d_alpha, d_beta};
        launch_fb(ctx, ctx->dev_meta.as<char>(), m.nb, ctx->dev_ck.as<float2>(), d_pi, d_pj, P, flags, o, 0, P,
                  static_cast<hipStream_t>(stream));
    });
}

int gsmc_malloc(void** ptr, uint64_t bytes)
{
    return guarded([&] { HIP_OK(hipMalloc(ptr, std::max<uint64_t>(bytes, 1))); });
}
int gsmc_free(void* ptr)
{
    return guarded([&] { HIP_OK(hipFree(ptr)); });
}
int gsmc_memcpy_htod(void* dst, const void* src, uint64_t bytes)
{
    return guarded([&] { HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int gsmc_memcpy_dtoh(void* dst, const void* src, uint64_t bytes)
{
    return guarded([&] { HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int gsmc_memset(void* dst, int value, uint64_t bytes)
{
    return guarded([&] { HIP_OK(hipMemset(dst, value, bytes)); });
}
int gsmc_synchronize(void)
{
    return guarded([&] { HIP_OK(hipDeviceSynchronize()); });
}
int gsmc_stream_create(void** stream)
{
    return guarded([&] {
        hipStream_t s;
        HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        *stream = s;
    });
}
int gsmc_stream_destroy(void* stream)
{
    return guarded([&] { HIP_OK(hipStreamDestroy(static_cast<hipStream_t>(stream))); });
}
int gsmc_stream_synchronize(void* stream)
{
    return guarded([&] { HIP_OK(hipStreamSynchronize(static_cast<hipStream_t>(stream))); });
}
int gsmc_event_create(void** event)
{
    return guarded([&] {
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        *event = e;
    });
}
int gsmc_event_destroy(void* event)
{
    return guarded([&] { HIP_OK(hipEventDestroy(static_cast<hipEvent_t>(event))); });
}
int gsmc_event_record(void* event, void* stream)
{
    return guarded([&] { HIP_OK(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream))); });
}
int gsmc_event_elapsed_ms(void* start, void* stop, float* ms)
{
    return guarded([&] {
        HIP_OK(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
        HIP_OK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    });
}

}  // extern "C"
