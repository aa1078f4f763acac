// Runtime and C ABI of the gfx950 Gamma-SMC decoder (include/gamma_smc_amd.h).
//
// A context owns, on one device: the bitpacked haplotype rows, the per-site
// gap descriptors, a shared reference to the transition tables, and
// reusable scratch (checkpoints, launch metadata, output slots, pinned
// staging). Tables are shared between contexts with the same
// (device, flow field, steps, scaled rates, precision) and freed with the
// last context that uses them.
#include <hip/hip_runtime.h>
#include <sys/mman.h>

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
#include "gene_kernels.h"
#include "tables.h"

// Device scratch of gsmc_run_fb_device_noalloc: launch metadata, flank
// states and checkpoints of one call, for callers that cannot allocate.
// Each device holds its own copy for the life of the process; host code
// uses the per-device address from hipGetSymbolAddress, never this name.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

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

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy at its own address).
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_at;

void* scratch_address(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_at.find(device);
    if (it != g_scratch_at.end()) return it->second;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        throw Error(GSMC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_at[device] = ptr;
    return ptr;
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
    // quiesced: the caller has already waited for every queued use of the
    // buffer (a workspace's buffers, after its completion event), so growing
    // it needs no device-wide wait.
    void ensure(size_t n, int dev, bool quiesced = false)
    {
        if (n <= bytes && p) return;
        // Growing a buffer that queued work may still read: let that work finish.
        if (p && !quiesced) HIP_OK(hipDeviceSynchronize());
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
    uint4* h2q = nullptr;   // forward-only path: cache[k] as FP16 quad entries, [N][51 * 50]
    ~DeviceTables()
    {
        if (base || h2q) (void)hipSetDevice(device);
        if (base) (void)hipFree(base);
        if (h2q) (void)hipFree(h2q);
    }
};

std::mutex g_tab_mu;
std::map<std::string, std::weak_ptr<DeviceTables>> g_tabs;
// The most recently used table set of each device stays alive after its
// last context closes, as the reference's process-wide table cache does
// (bindings.cpp:2016-2050, rebuilt only when the parameters change; there
// one set per process, here one per device), so back-to-back contexts and
// one-shot calls on the same parameters skip the build. Never destroyed:
// the process's end frees the device memory (a static destructor would run
// after the HIP runtime's).
auto* g_tabs_last = new std::map<int, std::shared_ptr<DeviceTables>>();

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
        if (auto sp = it->second.lock()) {
            (*g_tabs_last)[device] = sp;
            return sp;
        }
    }
    HostTables host;
    build_tables(ff, N, srho, smu, srho64, smu64, precision, host);
    auto dt = std::make_shared<DeviceTables>();
    dt->device = device;
    dt->n_steps = N;
    const size_t layers = (size_t)N * kLayerBytes;
    dt->bytes = 2 * layers + kSlabBytes;
    HIP_OK(hipMalloc(&dt->base, dt->bytes));
    float2* node_tmp = nullptr;
    {
        // Node-major (m, c) slabs (fwd, bwd, stretch: the cell slabs' order)
        // go up as built and are expanded into 32-byte cells on the device.
        const size_t n_slabs = (size_t)N * 4 + 1, node_slab = (size_t)kGridNodes * sizeof(float2);
        HIP_OK(hipMalloc(&node_tmp, n_slabs * node_slab));
        char* t = reinterpret_cast<char*>(node_tmp);
        HIP_OK(hipMemcpy(t, host.fwd.data(), (size_t)N * 2 * node_slab, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(t + (size_t)N * 2 * node_slab, host.bwd.data(), (size_t)N * 2 * node_slab,
                         hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(t + (size_t)N * 4 * node_slab, host.stretch_last.data(), node_slab, hipMemcpyHostToDevice));
        const int64_t entries = (int64_t)n_slabs * (int64_t)(kSlabBytes / 16);
        hipLaunchKernelGGL(cells_from_nodes_kernel, dim3((unsigned)((entries + 255) / 256)), dim3(256), 0, nullptr,
                           node_tmp, reinterpret_cast<float4*>(dt->base), (int64_t)n_slabs);
        HIP_OK(hipGetLastError());
    }
    // the kernels add offsets biased by kFloatIntBias (fb_kernels.h::lookup)
    const uintptr_t b0 = reinterpret_cast<uintptr_t>(dt->base) - kFloatIntBias;
    dt->view.fwd = reinterpret_cast<const char*>(b0);
    dt->view.bwd = reinterpret_cast<const char*>(b0 + layers);
    dt->view.stretch = reinterpret_cast<const char*>(b0 + 2 * layers);
    {
        // FP16 copy of the composed transitions, rounded on the device as
        // the reference rounds them on the host (__floats2half2_rn), laid
        // out as quad entries (one 16-byte gather per lookup).
        const int64_t nodes = (int64_t)N * kGridNodes;
        const int64_t quads = (int64_t)N * kQuadSlabEntries;
        float2* tmp = nullptr;
        HIP_OK(hipMalloc(&tmp, (size_t)nodes * sizeof(float2)));
        HIP_OK(hipMalloc(&dt->h2q, (size_t)quads * sizeof(uint4)));
        HIP_OK(hipMemcpy(tmp, host.cache.data(), (size_t)nodes * sizeof(float2), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(to_half2_quads_kernel, dim3((unsigned)((quads + 255) / 256)), dim3(256), 0, nullptr, tmp,
                           dt->h2q, (int64_t)N);
        HIP_OK(hipGetLastError());
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipFree(tmp));
        HIP_OK(hipFree(node_tmp));
    }
    g_tabs[key] = dt;
    (*g_tabs_last)[device] = dt;
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

// (int)(x) as the device computes it: truncate, saturating, NaN -> 0.
int trunc_steps(double x)
{
    if (x != x) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return INT32_MIN;
    return (int)x;
}

// Prior-after-stretches table bound: 4M entries (32 MB), i.e. positions up
// to 4e9 bp at the default 1000-step tables.
constexpr int64_t kMaxPriorStretches = 1 << 22;

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

struct Tuning {
    int K = 8;               // checkpoint interval = recomputed window; 1: every forward state stored
    int np = 1;              // pairs per lane (mean-only kernel; 1 or 2)
    int schedule = 0;        // 0 static, 1 persistent queue
    int xcd_order = 0;       // static schedule: give each XCD a contiguous run of items
    int64_t min_waves = 8192;
    uint64_t max_state_bytes = 0;  // cap on a launch's checkpoint / stored-state buffer (0: free memory)
};

struct gsmc_batch;
struct HostSlot;

struct gsmc_context {
    std::mutex mu;
    int device = 0;
    int cus = 0;
    int64_t n = 0, S = 0, W = 0;
    float Ne = 0.0f;
    int N = 1000;
    Tuning tune;
    std::vector<double> pos;
    DevBuf packed, desc;
    DevBuf prior;              // prior[q]: the prior after q full stretches (sequence starts)
    int64_t prior_n = 0;
    DevBuf desc_h2;            // forward-only path: per-site gap descriptors
    DevBuf fwd_buf;            // forward-only host path: pair chunk and outputs
    std::shared_ptr<DeviceTables> tabs;
    // host-output path: output slots, each a decode stream with its own
    // workspace and device rows, drained through pinned staging on `copy`
    DevBuf pairs;
    std::vector<std::unique_ptr<HostSlot>> slots;
    PinnedBuf stage[2];
    hipStream_t compute = nullptr, copy = nullptr;
    hipEvent_t ev_stage[2] = {nullptr, nullptr};   // staging buffer filled (copy -> host)
    std::unique_ptr<gsmc_batch> dev_batch;
    DevBuf sum_pairs, sum_site;  // per-site summary path: pair chunk, outputs and running values
    DevBuf gene_pairs, gene_part, gene_meta, gene_acc, gene_split;  // per-gene statistics path
    void* scratch = nullptr;     // this device's `scratch` (gsmc_run_fb_device_noalloc)

    ~gsmc_context();
};

// A launch workspace: metadata, checkpoints and queue counter of one launch
// in flight. A new launch on the workspace first waits for the previous one
// unless it repeats the same metadata on the same stream.
struct gsmc_batch {
    gsmc_context* ctx = nullptr;
    DevBuf meta, ck, queue, yb, part;
    PinnedBuf stage;
    std::vector<char> last;
    hipStream_t last_stream = nullptr;
    hipEvent_t done = nullptr;
    bool pending = false;
    // gsmc_batch_reserve sized the buffers: launches keep their state
    // buffer within the reservation (more phases, or fewer resident
    // waves) instead of growing it on the launch path.
    bool reserved = false;
    ~gsmc_batch()
    {
        if (done) {
            (void)hipSetDevice(ctx->device);
            if (pending) (void)hipEventSynchronize(done);
            (void)hipEventDestroy(done);
        }
    }
};

// One output slot of the host-output path: a decode stream, its launch
// workspace and device rows, and the events that hand the rows to the copy
// stream (kernel) and back (free).
struct HostSlot {
    std::unique_ptr<gsmc_batch> batch;
    DevBuf out;
    hipStream_t stream = nullptr;
    hipEvent_t ev_kernel = nullptr;  // rows written (slot stream -> copy)
    hipEvent_t ev_free = nullptr;    // rows drained (copy -> slot stream)
    int device = 0;
    ~HostSlot()
    {
        (void)hipSetDevice(device);
        batch.reset();
        if (stream) (void)hipStreamSynchronize(stream);
        if (ev_kernel) (void)hipEventDestroy(ev_kernel);
        if (ev_free) (void)hipEventDestroy(ev_free);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

gsmc_context::~gsmc_context()
{
    (void)hipSetDevice(device);
    dev_batch.reset();
    slots.clear();
    for (auto& e : ev_stage)
        if (e) (void)hipEventDestroy(e);
    if (compute) (void)hipStreamDestroy(compute);
    if (copy) (void)hipStreamDestroy(copy);
}

namespace gsmc {
namespace {

void init_streams(gsmc_context* ctx)
{
    if (ctx->compute) return;
    HIP_OK(hipStreamCreateWithFlags(&ctx->compute, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&ctx->copy, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) HIP_OK(hipEventCreateWithFlags(&ctx->ev_stage[i], hipEventDisableTiming));
}

std::unique_ptr<gsmc_batch> new_batch(gsmc_context* ctx)
{
    auto b = std::make_unique<gsmc_batch>();
    b->ctx = ctx;
    HIP_OK(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    return b;
}

// At least n output slots of the host-output path.
void ensure_slots(gsmc_context* ctx, int n)
{
    init_streams(ctx);
    while ((int)ctx->slots.size() < n) {
        auto sl = std::make_unique<HostSlot>();
        sl->device = ctx->device;
        HIP_OK(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&sl->ev_kernel, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&sl->ev_free, hipEventDisableTiming));
        sl->batch = new_batch(ctx);
        ctx->slots.push_back(std::move(sl));
    }
}

void finish_create(gsmc_context* ctx, const double* positions, double Ne, double mu, double rho, const char* path,
                   int cache_steps, int precision)
{
    ctx->N = cache_steps <= 0 ? 1000 : cache_steps;
    ctx->Ne = (float)Ne;
    ctx->pos.assign(positions, positions + ctx->S);
    ctx->tabs = get_tables(ctx->device, path, ctx->N, Ne, mu, rho, precision);
    // 64 zero entries past the last site: kernels load descriptors in
    // fixed-size runs.
    std::vector<uint32_t> d((size_t)ctx->S + 64, 0u);
    for (int64_t s = 0; s < ctx->S; ++s) {
        const double delta = s == 0 ? positions[0] + 1.0 : positions[s] - positions[s - 1];
        d[s] = descriptor(round_steps(delta), ctx->N);
    }
    ctx->desc.ensure(d.size() * 4, ctx->device);
    HIP_OK(hipMemcpy(ctx->desc.p, d.data(), d.size() * 4, hipMemcpyHostToDevice));
    // Forward-only path (gamma_smc_flow.cu:1212-1217): gap = (int)(delta + 0.5)
    // from the previous site, nothing at site 0 or for gaps that round to 0.
    for (int64_t s = 0; s < ctx->S; ++s) {
        const int gap = s == 0 ? 0 : trunc_steps(positions[s] - positions[s - 1] + 0.5);
        d[s] = gap > 0 ? descriptor(gap, ctx->N) : kNoAdvance;
    }
    ctx->desc_h2.ensure(d.size() * 4, ctx->device);
    HIP_OK(hipMemcpy(ctx->desc_h2.p, d.data(), d.size() * 4, hipMemcpyHostToDevice));
    // A sequence may start at any site; its first segment applies
    // (steps - 1) / N full stretches to the prior. Precompute the prior after
    // q stretches for every q any site can need (bounded; beyond the bound
    // the kernels apply the stretches themselves).
    int64_t qmax = 0;
    for (int64_t s = 0; s < ctx->S; ++s) qmax = std::max<int64_t>(qmax, (round_steps(positions[s] + 1.0) - 1) / ctx->N);
    ctx->prior_n = std::min<int64_t>(qmax, kMaxPriorStretches) + 1;
    ctx->prior.ensure((size_t)ctx->prior_n * sizeof(float2), ctx->device);
    hipLaunchKernelGGL(prior_table_kernel, dim3(1), dim3(64), 0, nullptr, ctx->tabs->view, ctx->prior.as<float2>(),
                       ctx->prior_n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, ctx->device));
    ctx->cus = prop.multiProcessorCount;
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
    float* site_mean;  // summary launches: per-site reductions over the pairs
    float* site_min;
    float* site_max;
};

// One job of a launch: pairs (device ids) decoded over `blocks` (global site
// windows) of the sequence that starts at site seq_lo; outputs are
// [rows][P] with row r = site row0 + r.
struct JobSpec {
    int64_t seq_lo = 0;
    std::vector<Block> blocks;
    const int32_t* pi = nullptr;
    const int32_t* pj = nullptr;
    int64_t P = 0;
    Outputs out{};
    int64_t row0 = 0;
    // Summary launches over pair chunks: the first chunk starts the running
    // per-site values, the last one writes the outputs (over P_total pairs).
    bool sum_first = true, sum_last = true;
    double* tot = nullptr;
    float* tmin = nullptr;
    float* tmax = nullptr;
    int64_t P_total = 0;
    // Per-gene launches: piece of each site, piece sums out
    const int32_t* gene_piece = nullptr;
    double2* gene_part = nullptr;
};

// Summary partials of one job in a launch.
struct SumJob {
    const JobSpec* spec;
    int64_t part_off, groups, rows;
};

// Launch metadata: [jobs][blocks][items], 256-byte aligned sections.
// A run of items launched as one fwd / bwd kernel pair. Phases run one
// after the other and reuse the state buffer.
struct Phase {
    int64_t first = 0, count = 0;
};

struct Plan {
    std::vector<char> bytes;
    size_t off_blocks = 0, off_items = 0;
    int64_t n_items = 0;
    int64_t n_waves = 0;  // waves with work (items without padding)
    size_t ck_elems = 0;  // state buffer (float2) of the largest phase
    size_t part_elems = 0;        // summary partials (per plane)
    std::vector<SumJob> sums;
    std::vector<Phase> phases;
    int64_t fused_grid = 0;   // stored states: resident waves of the fused kernel
    int64_t slot_elems = 0;   // ... and the float2 state elements of each wave's slot
    int np = 1;               // pairs per lane of the plan's tiles (kernels index 64 np lanes per item)
};

// Most pairs per lane any kernel is built for (sizes of reserved buffers).
constexpr int kMaxPairsPerLane = 4;

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Forward states stored at every core site (instead of checkpoints every K
// sites and a recompute): tuning K = 1, for the mean / code / summary
// kernels. The forward walk goes in windows of 16 sites and the backward
// walk in windows of kStoredBwdK (short windows keep it at six waves per
// SIMD): as two kernels per phase of items (static schedule), or as one
// fused persistent kernel with a state slot per resident wave (persistent
// schedule).
// Internal launch flag of per-gene statistics: the backward kernel
// accumulates per-gene piece sums instead of storing rows (one pair per
// lane, checkpoint windows, static schedule).
constexpr int kGeneLaunch = 1 << 8;

// Flags a caller may pass (the GSMC_* run flags); other bits, the internal
// launch flags among them, are refused.
constexpr int kPublicFlags = GSMC_CI | GSMC_POSTERIOR | GSMC_SUMMARY | GSMC_Q8 | GSMC_Q4;
void check_flags(int flags)
{
    require((flags & ~kPublicFlags) == 0, "unknown flag bits: " + std::to_string(flags & ~kPublicFlags));
}

bool stored_states(const Tuning& tu, int flags)
{
    return tu.K == 1 && !(flags & (GSMC_CI | GSMC_POSTERIOR | kGeneLaunch));
}
int window_k(const Tuning& tu) { return tu.K == 8 ? 8 : 16; }

// The tuning one launch runs with: the context's, except that a per-site
// summary over checkpoints of 8 sites runs the static schedule. Its
// persistent backward kernel needs 96 registers (5 waves per SIMD) where the
// static one needs 80 (6 waves), and measured 6% slower (134 against 126 ms
// for the chromosome workload on one box).
Tuning launch_tuning(const Tuning& t, int flags)
{
    Tuning u = t;
    if ((flags & GSMC_SUMMARY) && !stored_states(t, flags) && t.K == 8 && t.np == 1) u.schedule = 0;
    if (flags & kGeneLaunch) u.schedule = 0;
    return u;
}
constexpr int kStoredBwdK = 4;
constexpr int kStoredWaves = 6;

const void* fused_kernel(int flags)
{
    return (flags & GSMC_SUMMARY) ? reinterpret_cast<const void*>(&fb_fused_queue<16, kStoredBwdK, 1, true, kStoredWaves>)
                                  : reinterpret_cast<const void*>(&fb_fused_queue<16, kStoredBwdK, 1, false, kStoredWaves>);
}
int64_t resident(const gsmc_context* ctx, const void* fn);

// Bytes the state buffer of workspace b may take for a plan needing `want`:
// all of it when that fits in the current allocation or in the allocation
// plus free device memory (less a reserve), else that much. A reserved
// workspace (fixed) never grows: the plan fits in what it holds.
size_t state_budget(const gsmc_context* ctx, size_t have, size_t want, bool fixed)
{
    const size_t cap = ctx->tune.max_state_bytes ? (size_t)ctx->tune.max_state_bytes : SIZE_MAX;
    if (want <= have && want <= cap) return want;
    if (fixed) return std::min(std::min(want, have), cap);
    size_t free_b = 0, total_b = 0;
    HIP_OK(hipMemGetInfo(&free_b, &total_b));
    const size_t reserve = (size_t)2 << 30;
    const size_t avail = have + (free_b > reserve ? free_b - reserve : 0);
    return std::min(std::min(want, avail), cap);
}

Plan make_plan(const gsmc_context* ctx, const std::vector<JobSpec>& jobs, const Tuning& tu, int np,
               bool summary, bool stored, size_t have_ck_bytes, int flags, bool fixed_budget = false)
{
    std::vector<JobDev> jd;
    std::vector<BlockDev> bd;
    std::vector<ItemDev> items;
    std::vector<SumJob> sums;
    size_t ck = 0, part = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
        const JobSpec& js = jobs[j];
        if (js.P == 0 || js.blocks.empty()) continue;
        const int64_t tile_pairs = 64 * (int64_t)np;
        const int64_t tiles = (js.P + tile_pairs - 1) / tile_pairs;
        const int64_t rows = (int64_t)js.blocks.back().ce - js.row0;
        int64_t part_off = -1;
        if (summary) {
            part_off = (int64_t)part;
            sums.push_back({&js, part_off, tiles * np, rows});
            part += (size_t)(tiles * np) * (size_t)rows;
        }
        jd.push_back({js.out.mean, js.out.lower, js.out.upper, js.out.alpha, js.out.beta, js.pi, js.pj, js.P,
                      js.row0, rows, part_off});
        for (const Block& b : js.blocks) {
            BlockDev d;
            d.cs = b.cs;
            d.ce = b.ce;
            d.ps = b.ps;
            d.pe = b.pe;
            d.q0 = 0;
            d.pad = 0;
            if (b.ps == js.seq_lo) {
                // first site of the decoded sequence: its gap from position -1.
                // The full stretches of that gap start from the prior, so the
                // walk starts at prior[q] and applies only the last layer.
                // The backward walk's step at this site comes after the
                // site's combine and is never used: one lookup stands in.
                const uint32_t full = descriptor(round_steps(ctx->pos[b.ps] + 1.0), ctx->N);
                const int64_t q = full >> kDescLayerBits;
                if (q < ctx->prior_n) {
                    d.sd_fwd = full & kDescLayerMask;
                    d.q0 = (int32_t)q;
                } else {
                    d.sd_fwd = full;
                }
                d.sd_bwd = 0;
            } else {
                // block restart: one bp from (position - 1), as the block kernels do
                const double pp = ctx->pos[b.ps];
                d.sd_fwd = descriptor(round_steps(pp - (pp - 1.0)), ctx->N);
                d.sd_bwd = 0;
            }
            d.job = (int32_t)(jd.size() - 1);
            const int kw = stored ? kStoredBwdK : window_k(tu);
            d.nwin = (int32_t)((b.ce - b.cs + kw - 1) / kw);
            // float2 rows: two per 16-byte record of two checkpoint windows, or of two
            // core sites (stored states)
            const int64_t rows = stored ? 2 * (int64_t)((b.ce - b.cs + 1) / 2) : 2 * (int64_t)((d.nwin + 1) / 2);
            require(bd.size() < (size_t)INT32_MAX, "too many block windows");
            // ck_off holds the item's state elements until phases place them
            for (int64_t t = 0; t < tiles; ++t) items.push_back({(int32_t)bd.size(), (int32_t)t, rows * tile_pairs});
            ck += (size_t)(tiles * rows * tile_pairs);
            bd.push_back(d);
        }
    }
    Plan pl;
    pl.np = np;
    pl.n_waves = (int64_t)items.size();
    // Longest chains first: an item's duration is its walk length (forward
    // over [ps, ce), backward over the right flank, two lookups per core
    // site), so ragged batches dispatch their long items before the short
    // ones that fill in around them. Items of equal length keep their order.
    std::stable_sort(items.begin(), items.end(), [&](const ItemDev& x, const ItemDev& y) {
        auto cost = [&](const ItemDev& it) {
            const BlockDev& b = bd[(size_t)it.block];
            return (int64_t)(b.pe - b.ps) + 2 * (int64_t)(b.ce - b.cs);
        };
        return cost(x) > cost(y);
    });
    if (stored && tu.schedule == 1) {
        // one fused launch: a state slot (the largest item's rows) per
        // resident wave, as many waves as fit in the budget
        size_t slot = 0;
        for (const ItemDev& it : items) slot = std::max(slot, (size_t)it.ck_off);
        slot = std::max<size_t>(slot, 1);
        const int64_t want = std::max<int64_t>(1, std::min<int64_t>(resident(ctx, fused_kernel(flags)),
                                                                    (int64_t)items.size()));
        const size_t budget =
            state_budget(ctx, have_ck_bytes, (size_t)want * slot * sizeof(float2), fixed_budget) / sizeof(float2);
        if (slot > budget)
            throw Error(GSMC_ENOMEM, "state budget below the forward states of one work item (" +
                                         std::to_string(slot * sizeof(float2)) + " bytes)");
        pl.fused_grid = std::min<int64_t>(want, (int64_t)(budget / slot));
        pl.slot_elems = (int64_t)slot;
        for (ItemDev& it : items) it.ck_off = 0;
        pl.phases.push_back({0, (int64_t)items.size()});
        pl.n_items = (int64_t)items.size();
        pl.ck_elems = (size_t)pl.fused_grid * slot;
        pl.part_elems = part;
        pl.sums = std::move(sums);
        const size_t sj = jd.size() * sizeof(JobDev), sb = bd.size() * sizeof(BlockDev);
        pl.off_blocks = align256(sj);
        pl.off_items = pl.off_blocks + align256(sb);
        pl.bytes.assign(pl.off_items + items.size() * sizeof(ItemDev), 0);
        if (sj) std::memcpy(pl.bytes.data(), jd.data(), sj);
        if (sb) std::memcpy(pl.bytes.data() + pl.off_blocks, bd.data(), sb);
        if (!items.empty())
            std::memcpy(pl.bytes.data() + pl.off_items, items.data(), items.size() * sizeof(ItemDev));
        return pl;
    }
    // Phases: when the states of all items exceed the budget, deal the
    // items round-robin (longest first) into the fewest phases that fit, so
    // every phase gets the same mix of long and short items.
    const size_t budget = state_budget(ctx, have_ck_bytes, ck * sizeof(float2), fixed_budget) / sizeof(float2);
    int64_t n_ph = 1;
    if (ck > budget) {
        size_t biggest = 0;
        for (const ItemDev& it : items) biggest = std::max(biggest, (size_t)it.ck_off);
        if (biggest > budget)
            throw Error(GSMC_ENOMEM, "state budget below the checkpoints / forward states of one work item (" +
                                         std::to_string(biggest * sizeof(float2)) + " bytes)");
        n_ph = (int64_t)((ck + budget - 1) / budget);
    }
    std::vector<std::vector<ItemDev>> ph;
    for (;; ++n_ph) {
        ph.assign((size_t)n_ph, {});
        std::vector<size_t> used((size_t)n_ph, 0);
        for (size_t i = 0; i < items.size(); ++i) {
            ph[i % (size_t)n_ph].push_back(items[i]);
            used[i % (size_t)n_ph] += (size_t)items[i].ck_off;
        }
        if (*std::max_element(used.begin(), used.end()) <= budget) break;
    }
    std::vector<ItemDev> all;
    size_t ck_max = 0;
    for (std::vector<ItemDev>& v : ph) {
        if (tu.schedule == 0 && tu.xcd_order && !v.empty()) {
            // Workgroup g runs on XCD g % 8: hand each XCD a contiguous run of
            // items, so the waves sharing an XCD's L2 walk the same blocks
            // (and so the same table layers) at the same time.
            const int64_t G = (int64_t)v.size();
            const int64_t per = (G + 7) / 8;
            std::vector<ItemDev> perm((size_t)(per * 8), ItemDev{-1, 0, 0});
            for (int64_t g = 0; g < per * 8; ++g) {
                const int64_t src = (g % 8) * per + g / 8;
                if (src < G) perm[(size_t)g] = v[(size_t)src];
            }
            v.swap(perm);
        }
        size_t off = 0;
        for (ItemDev& it : v) {
            const size_t n = it.block >= 0 ? (size_t)it.ck_off : 0;
            it.ck_off = (int64_t)off;
            off += n;
        }
        ck_max = std::max(ck_max, off);
        pl.phases.push_back({(int64_t)all.size(), (int64_t)v.size()});
        all.insert(all.end(), v.begin(), v.end());
    }
    items.swap(all);
    pl.n_items = (int64_t)items.size();
    pl.ck_elems = ck_max;
    pl.part_elems = part;
    pl.sums = std::move(sums);
    const size_t sj = jd.size() * sizeof(JobDev), sb = bd.size() * sizeof(BlockDev);
    pl.off_blocks = align256(sj);
    pl.off_items = pl.off_blocks + align256(sb);
    pl.bytes.assign(pl.off_items + items.size() * sizeof(ItemDev), 0);
    if (sj) std::memcpy(pl.bytes.data(), jd.data(), sj);
    if (sb) std::memcpy(pl.bytes.data() + pl.off_blocks, bd.data(), sb);
    if (!items.empty()) std::memcpy(pl.bytes.data() + pl.off_items, items.data(), items.size() * sizeof(ItemDev));
    return pl;
}

// Pairs per lane a launch uses: the mean-only kernels (the throughput path)
// are built for 1 and 2, the CI / posterior kernels for 1.
int launch_np(const Tuning& tu, int flags)
{
    if (flags & (GSMC_CI | GSMC_POSTERIOR | kGeneLaunch)) return 1;
    // stored forward states: one or two pairs per lane with the static
    // schedule, one in the fused persistent kernel
    if (tu.K == 1) return tu.schedule == 1 ? 1 : std::min(tu.np, 2);
    // four pairs per lane: mean / code kernels only; summaries take two
    return (flags & GSMC_SUMMARY) ? std::min(tu.np, 2) : tu.np;
}

// Kernel pair of one launch shape, and each kernel's resident waves.
struct KernelPair {
    const void* fwd_items;
    const void* bwd_items;
    const void* fwd_queue;
    const void* bwd_queue;
    const void* fwd_grab;
    const void* bwd_grab;
};

template <int K, int NP, bool C, bool Q, bool SUM = false, bool ST = false> KernelPair kernels_of()
{
    return {reinterpret_cast<const void*>(&fb_fwd_items<K, NP, ST>),
            reinterpret_cast<const void*>(&fb_bwd_items<K, NP, C, Q, SUM, ST>),
            reinterpret_cast<const void*>(&fb_fwd_queue<K, NP, ST>),
            reinterpret_cast<const void*>(&fb_bwd_queue<K, NP, C, Q, SUM, ST>),
            reinterpret_cast<const void*>(&fb_fwd_grab<K, NP, ST>),
            reinterpret_cast<const void*>(&fb_bwd_grab<K, NP, C, Q, SUM, ST>)};
}

// The kernels for the pairs per lane launch_np plans with (the tile size of
// the plan and the kernel's NP must agree).
template <int K> KernelPair kernels_k(const Tuning& tu, int flags)
{
    const bool ci = flags & GSMC_CI, post = flags & GSMC_POSTERIOR;
    const int np = launch_np(tu, flags);
    if (flags & kGeneLaunch) {
        // static schedule only (launch_tuning): the items kernels fill every slot
        const void* f = reinterpret_cast<const void*>(&fb_fwd_items<K, 1, false>);
        const void* g = reinterpret_cast<const void*>(&fb_bwd_genes<K>);
        return {f, g, f, g, f, g};
    }
    if (flags & GSMC_SUMMARY) return np == 2 ? kernels_of<K, 2, false, false, true>() : kernels_of<K, 1, false, false, true>();
    if (ci && post) return kernels_of<K, 1, true, true>();
    if (ci) return kernels_of<K, 1, true, false>();
    if (post) return kernels_of<K, 1, false, true>();
    if constexpr (K == 8)
        if (np == 4) return kernels_of<K, 4, false, false>();
    if (np == 2) return kernels_of<K, 2, false, false>();
    return kernels_of<K, 1, false, false>();
}

template <int NP, bool SUM> KernelPair stored_kernels()
{
    return {reinterpret_cast<const void*>(&fb_fwd_items<16, NP, true>),
            reinterpret_cast<const void*>(&fb_bwd_items<kStoredBwdK, NP, false, false, SUM, true>),
            reinterpret_cast<const void*>(&fb_fwd_queue<16, NP, true>),
            reinterpret_cast<const void*>(&fb_bwd_queue<kStoredBwdK, NP, false, false, SUM, true>),
            reinterpret_cast<const void*>(&fb_fwd_grab<16, NP, true>),
            reinterpret_cast<const void*>(&fb_bwd_grab<kStoredBwdK, NP, false, false, SUM, true>)};
}

KernelPair kernels_for(const Tuning& tu, int flags)
{
    if (stored_states(tu, flags)) {
        const bool sum = (flags & GSMC_SUMMARY) != 0;
        if (launch_np(tu, flags) == 2) return sum ? stored_kernels<2, true>() : stored_kernels<2, false>();
        return sum ? stored_kernels<1, true>() : stored_kernels<1, false>();
    }
    return window_k(tu) == 8 ? kernels_k<8>(tu, flags) : kernels_k<16>(tu, flags);
}

int64_t resident(const gsmc_context* ctx, const void* fn)
{
    int per_cu = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, 0));
    return (int64_t)std::max(per_cu, 1) * ctx->cus;
}

void launch(const void* fn, unsigned grid, const FbArgs& a, hipStream_t st)
{
    void* args[] = {const_cast<FbArgs*>(&a)};
    HIP_OK(hipLaunchKernel(fn, dim3(grid), dim3(64), args, 0, st));
}

// Flank states of n_items items of 64 np pairs each (one float2 per pair).
size_t yb_bytes(int64_t n_items, int np) { return (size_t)std::max<int64_t>(n_items, 1) * 64 * (size_t)np * sizeof(float2); }

// Buffers a planned launch needs on workspace b. Every caller has waited
// for the workspace's previous launch (or the workspace is new), so no
// queued work reads these buffers and growing them needs no device-wide
// wait: a launch on one workspace never stalls on another's kernels.
void ensure_buffers(gsmc_batch* b, const Plan& pl, bool summary)
{
    gsmc_context* ctx = b->ctx;
    b->meta.ensure(pl.bytes.size(), ctx->device, true);
    b->ck.ensure(std::max<size_t>(pl.ck_elems, 1) * sizeof(float2), ctx->device, true);
    b->queue.ensure(256, ctx->device, true);
    b->yb.ensure(yb_bytes(pl.n_items, pl.np), ctx->device, true);
    if (summary)
        b->part.ensure(std::max<size_t>(pl.part_elems, 1) * (sizeof(double) + 2 * sizeof(float)), ctx->device, true);
}

// Kernel arguments of a planned launch on workspace b's buffers.
FbArgs make_args(gsmc_batch* b, const Plan& pl, int flags)
{
    gsmc_context* ctx = b->ctx;
    // the kernels index these buffers by the plan: refuse a launch they cannot hold
    require(b->yb.bytes >= yb_bytes(pl.n_items, pl.np) && b->ck.bytes >= pl.ck_elems * sizeof(float2),
            "internal: launch buffers smaller than the plan");
    FbArgs a;
    a.packed = ctx->packed.as<uint64_t>();
    a.n_haps = ctx->n;
    a.W = ctx->W;
    a.desc = ctx->desc.as<uint32_t>();
    a.prior = ctx->prior.as<float2>();
    a.qbits = (flags & GSMC_Q8) ? 8 : (flags & GSMC_Q4) ? 4 : 0;
    a.q_inv_range = 1.0f / logf(20.0f * ctx->Ne);  // codes span [log 1, log 20Ne] (bindings.cpp:1698-1700)
    a.jobs = reinterpret_cast<const JobDev*>(b->meta.p);
    a.blocks = reinterpret_cast<const BlockDev*>(b->meta.as<char>() + pl.off_blocks);
    a.items = reinterpret_cast<const ItemDev*>(b->meta.as<char>() + pl.off_items);
    a.n_items = pl.n_items;
    a.queue = b->queue.as<int>();
    a.queue2 = b->queue.as<int>() + 1;
    a.slot_elems = pl.slot_elems;
    a.ck = b->ck.as<float2>();
    a.yb = b->yb.as<float2>();
    a.t = ctx->tabs->view;
    a.unscale = 2.0f * ctx->Ne;
    a.psum = nullptr;
    a.pmin = a.pmax = nullptr;
    if (flags & GSMC_SUMMARY) {
        const size_t n = std::max<size_t>(pl.part_elems, 1);
        a.psum = b->part.as<double>();
        a.pmin = reinterpret_cast<float*>(a.psum + n);
        a.pmax = a.pmin + n;
    }
    return a;
}

// The kernels of a planned launch, in order: per phase the queue reset
// (persistent schedule) and the two decode kernels, then the summary
// epilogues. No copies and no memsets, so the sequence can also be
// captured into a graph.
void enqueue_kernels(gsmc_context* ctx, const Plan& pl, const FbArgs& a, int flags, hipStream_t stream)
{
    const Tuning tu = launch_tuning(ctx->tune, flags);
    if (pl.fused_grid > 0) {
        hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(1), 0, stream, a.queue);
        launch(fused_kernel(flags), (unsigned)pl.fused_grid, a, stream);
    }
    const KernelPair kp = kernels_for(tu, flags);
    for (const Phase& ph : pl.phases) {
        if (pl.fused_grid > 0) break;
        if (ph.count == 0) continue;
        FbArgs pa = a;
        pa.items = a.items + ph.first;
        pa.n_items = ph.count;
        if (tu.schedule == 1) {
            hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(1), 0, stream, a.queue);
            launch(kp.fwd_queue, (unsigned)std::max<int64_t>(1, std::min(resident(ctx, kp.fwd_queue), ph.count)), pa,
                   stream);
            launch(kp.bwd_queue, (unsigned)std::max<int64_t>(1, std::min(resident(ctx, kp.bwd_queue), ph.count)), pa,
                   stream);
        } else if (tu.schedule == 2) {
            // one workgroup per item, each taking the next item of the counter
            hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(1), 0, stream, a.queue);
            launch(kp.fwd_grab, (unsigned)ph.count, pa, stream);
            launch(kp.bwd_grab, (unsigned)ph.count, pa, stream);
        } else {
            launch(kp.fwd_items, (unsigned)ph.count, pa, stream);
            launch(kp.bwd_items, (unsigned)ph.count, pa, stream);
        }
    }
    for (const SumJob& sj : pl.sums) {
        const JobSpec& js = *sj.spec;
        const unsigned grid = (unsigned)((sj.rows + 255) / 256);
        hipLaunchKernelGGL(summary_reduce_kernel, dim3(grid), dim3(256), 0, stream, a.psum + sj.part_off,
                           a.pmin + sj.part_off, a.pmax + sj.part_off, sj.groups, sj.rows, js.tot, js.tmin, js.tmax,
                           (int)js.sum_first, (int)js.sum_last, (double)(js.P_total > 0 ? js.P_total : js.P),
                           js.out.site_mean, js.out.site_min, js.out.site_max);
    }
    HIP_OK(hipGetLastError());
}

// Queue one launch of `jobs` on `stream` using workspace b.
void launch_plan(gsmc_batch* b, const std::vector<JobSpec>& jobs, int flags, hipStream_t stream)
{
    gsmc_context* ctx = b->ctx;
    const Tuning tu = launch_tuning(ctx->tune, flags);
    const bool summary = (flags & GSMC_SUMMARY) != 0;
    Plan pl = make_plan(ctx, jobs, tu, launch_np(tu, flags), summary, stored_states(tu, flags), b->ck.bytes, flags,
                        b->reserved);
    if (pl.n_items == 0) return;
    const bool same = b->pending && stream == b->last_stream && pl.bytes == b->last;
    if (!same) {
        if (b->pending) HIP_OK(hipEventSynchronize(b->done));
        b->pending = false;
        ensure_buffers(b, pl, summary);
        b->stage.ensure(pl.bytes.size());
        std::memcpy(b->stage.p, pl.bytes.data(), pl.bytes.size());
        HIP_OK(hipMemcpyAsync(b->meta.p, b->stage.p, pl.bytes.size(), hipMemcpyHostToDevice, stream));
        b->last.swap(pl.bytes);
        b->last_stream = stream;
    }
    FbArgs a = make_args(b, pl, flags);
    if (flags & kGeneLaunch) {
        require(jobs.size() == 1 && jobs[0].gene_piece && jobs[0].gene_part, "internal: gene launch without pieces");
        a.gene_piece = jobs[0].gene_piece;
        a.gene_part = jobs[0].gene_part;
    }
    enqueue_kernels(ctx, pl, a, flags, stream);
    HIP_OK(hipEventRecord(b->done, stream));
    b->pending = true;
}

int n_outputs(int flags) { return 1 + ((flags & GSMC_CI) ? 2 : 0) + ((flags & GSMC_POSTERIOR) ? 2 : 0); }

// Output encodings: log-quantized codes replace the float mean only, and
// uint4 codes pack two rows per byte, so every block core must start at an
// even row of its job.
void check_codes(int flags, const std::vector<Block>& blocks, int64_t row0)
{
    const bool codes = (flags & (GSMC_Q8 | GSMC_Q4)) != 0;
    if (!codes) return;
    require(!((flags & GSMC_Q8) && (flags & GSMC_Q4)), "GSMC_Q8 and GSMC_Q4 are exclusive");
    require(!(flags & (GSMC_CI | GSMC_POSTERIOR | GSMC_SUMMARY)),
            "quantized output cannot be combined with GSMC_CI, GSMC_POSTERIOR or GSMC_SUMMARY");
    if (flags & GSMC_Q4)
        for (const Block& b : blocks)
            require(((int64_t)b.cs - row0) % 2 == 0, "uint4 output needs an even core_block_sites");
}

void check_pairs(const gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P)
{
    // the message is built only for a bad pair (building it for every pair
    // took 18 ms of a 499,500-pair call)
    for (int64_t p = 0; p < P; ++p)
        if (!(pi[p] >= 0 && pi[p] < ctx->n && pj[p] >= 0 && pj[p] < ctx->n))
            throw Error(GSMC_EINVAL,
                        "pair index out of range: (" + std::to_string(pi[p]) + ", " + std::to_string(pj[p]) + ")");
}

// Parallel host copy of staged rows into the caller's [S][P] arrays.
void scatter_rows(float* dst, int64_t ld_dst, int64_t col0, const float* src, int64_t rows, int64_t cols)
{
    // one thread per 16 MB, at most 16 (the host share of one GPU)
    const size_t bytes = (size_t)rows * cols * 4;
    const int nt = (int)std::min<size_t>(16, std::max<size_t>(1, bytes >> 24));
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
// groups of consecutive blocks, each group decoded by one launch into the
// rows of an output slot, then drained to host through pinned staging while
// the next groups run. n_slots slots (at least two) take the groups in
// turn; each slot decodes on its own stream with its own workspace, so
// launches of different slots may overlap on the device, and the copy
// stream drains the slots in group order.
void run_to_host(gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P, const std::vector<Block>& blocks,
                 int64_t pair_batch, int flags, float* const* host_out, int n_slots)
{
    require(!(flags & (GSMC_Q8 | GSMC_Q4 | GSMC_SUMMARY)),
            "host-output runs write float rows; use gsmc_run_fb_device, a batch launch or gsmc_run_fb_summary");
    n_slots = std::max(2, n_slots);
    ensure_slots(ctx, n_slots);
    const int nout = n_outputs(flags);
    const int64_t chunk = (pair_batch > 0) ? std::min<int64_t>(pair_batch, P) : P;
    // Blocks per launch: enough tile x block waves to fill the chip, bounded
    // by a device budget for the output slots (half of free memory, shared).
    size_t free_b = 0, total_b = 0;
    HIP_OK(hipMemGetInfo(&free_b, &total_b));
    size_t have_slots = 0;
    for (int s = 0; s < n_slots; ++s) have_slots += ctx->slots[(size_t)s]->out.bytes;
    const int64_t tile_pairs = 64 * (int64_t)launch_np(ctx->tune, flags);
    const int64_t tiles = (chunk + tile_pairs - 1) / tile_pairs;
    int64_t per_group = std::max<int64_t>(1, (ctx->tune.min_waves + tiles - 1) / tiles);
    int64_t max_rows = 0;
    for (const Block& b : blocks) max_rows = std::max<int64_t>(max_rows, b.ce - b.cs);
    const size_t slot_budget = (size_t)((double)(free_b + have_slots) * 0.5 / n_slots);
    while (per_group > 1 && (size_t)per_group * max_rows * chunk * 4 * nout > slot_budget) per_group /= 2;

    ctx->pairs.ensure((size_t)P * 8, ctx->device);
    int32_t* d_pi = ctx->pairs.as<int32_t>();
    int32_t* d_pj = d_pi + P;
    HIP_OK(hipMemcpy(d_pi, pi, (size_t)P * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_pj, pj, (size_t)P * 4, hipMemcpyHostToDevice));

    // The caller's output arrays are usually fresh, untouched allocations:
    // ask for huge pages, so the scatter's first touch of each page (a
    // page fault that zeroes it) happens per 2 MB instead of per 4 KB.
    // Advice only; nothing changes where it is not honoured.
    if (!blocks.empty()) {
        const size_t page = 4096;
        for (int o = 0; o < nout; ++o) {
            const uintptr_t lo = reinterpret_cast<uintptr_t>(host_out[o]);
            const uintptr_t hi = lo + (size_t)blocks.back().ce * (size_t)P * sizeof(float);
            const uintptr_t a0 = (lo + page - 1) & ~(uintptr_t)(page - 1), a1 = hi & ~(uintptr_t)(page - 1);
            if (a1 > a0) (void)madvise(reinterpret_cast<void*>(a0), a1 - a0, MADV_HUGEPAGE);
        }
    }

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
    for (int s = 0; s < n_slots; ++s) {
        size_t slot_max = 0;
        for (size_t i = (size_t)s; i < tasks.size(); i += (size_t)n_slots)
            slot_max = std::max(slot_max, (size_t)tasks[i].rows * tasks[i].np * 4 * nout);
        if (slot_max) ctx->slots[(size_t)s]->out.ensure(slot_max, ctx->device);
    }

    // A task holding every pair writes whole host rows: its planes go
    // straight into the caller's arrays when small (no staging buffers, no
    // host scatter: what small one-shot decodes spend most of their host
    // time on); larger ones stream through pinned staging, which the
    // runtime's copies from pageable memory do not match in bandwidth.
    constexpr size_t kDirectCopyMax = (size_t)64 << 20;
    auto direct = [&](const Task& t) { return t.np == P && (size_t)t.rows * t.np * 4 <= kDirectCopyMax; };
    bool need_stage = false;
    for (const Task& t : tasks) need_stage = need_stage || !direct(t);
    std::vector<char> slot_used((size_t)n_slots, 0);
    const size_t stage_bytes = (size_t)256 << 20;
    if (need_stage)
        for (int s = 0; s < 2; ++s) ctx->stage[s].ensure(stage_bytes);
    int stage_next = 0;
    bool stage_busy[2] = {false, false};
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

    // copy one finished slot out through the staging ring, in row pieces
    auto copy_out = [&](int slot, const Task& t) {
        HostSlot& sl = *ctx->slots[(size_t)slot];
        const size_t plane = (size_t)t.rows * t.np;
        const int64_t rows_per_piece = std::max<int64_t>(1, (int64_t)(stage_bytes / 4 / std::max<int64_t>(t.np, 1)));
        HIP_OK(hipStreamWaitEvent(ctx->copy, sl.ev_kernel, 0));
        if (direct(t)) {
            for (int o = 0; o < nout; ++o)
                HIP_OK(hipMemcpyAsync(host_out[o] + (size_t)t.row0 * P, sl.out.as<float>() + (size_t)o * plane,
                                      plane * 4, hipMemcpyDeviceToHost, ctx->copy));
            HIP_OK(hipEventRecord(sl.ev_free, ctx->copy));
            return;
        }
        for (int o = 0; o < nout; ++o) {
            const float* src = sl.out.as<float>() + (size_t)o * plane;
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
        HIP_OK(hipEventRecord(sl.ev_free, ctx->copy));
    };

    // Groups are launched n_slots - 1 ahead of the copies: group i's copy is
    // queued once group i + n_slots - 1 has been launched.
    const size_t lag = (size_t)n_slots - 1;
    for (size_t i = 0; i < tasks.size() + lag; ++i) {
        if (i < tasks.size()) {
            const Task& t = tasks[i];
            const int slot = (int)(i % (size_t)n_slots);
            HostSlot& sl = *ctx->slots[(size_t)slot];
            const size_t plane = (size_t)t.rows * t.np;
            if (slot_used[(size_t)slot]) HIP_OK(hipStreamWaitEvent(sl.stream, sl.ev_free, 0));
            slot_used[(size_t)slot] = 1;
            float* base = sl.out.as<float>();
            JobSpec js;
            js.seq_lo = 0;
            js.blocks.assign(blocks.begin() + t.b0, blocks.begin() + t.b0 + t.nb);
            js.pi = d_pi + t.p0;
            js.pj = d_pj + t.p0;
            js.P = t.np;
            js.row0 = t.row0;
            int k = 0;
            js.out.mean = base + plane * k++;
            if (flags & GSMC_CI) {
                js.out.lower = base + plane * k++;
                js.out.upper = base + plane * k++;
            }
            if (flags & GSMC_POSTERIOR) {
                js.out.alpha = base + plane * k++;
                js.out.beta = base + plane * k++;
            }
            launch_plan(sl.batch.get(), {js}, flags, sl.stream);
            HIP_OK(hipEventRecord(sl.ev_kernel, sl.stream));
        }
        if (i >= lag) copy_out((int)((i - lag) % (size_t)n_slots), tasks[i - lag]);
    }
    drain_stage(stage_next);
    drain_stage(stage_next ^ 1);
    for (int s = 0; s < n_slots; ++s) HIP_OK(hipStreamSynchronize(ctx->slots[(size_t)s]->stream));
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
        ctx->scratch = scratch_address(ctx->device);
        ctx->n = n;
        ctx->S = S;
        require(n <= INT32_MAX, "n_haps too large");
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

int gsmc_context_set_tuning(gsmc_context* ctx, int k, int pairs_per_lane, int schedule, int xcd_order,
                            int64_t min_waves)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        require(k == 1 || k == 8 || k == 16, "checkpoint_k must be 1 (stored forward states), 8 or 16");
        require(pairs_per_lane == 1 || pairs_per_lane == 2 || pairs_per_lane == 4, "pairs_per_lane must be 1, 2 or 4");
        require(pairs_per_lane != 4 || k == 8, "pairs_per_lane 4 needs checkpoint_k 8");
        require(schedule >= 0 && schedule <= 2, "schedule must be 0 (static), 1 (persistent) or 2 (dynamic)");
        require(min_waves >= 1, "min_waves_per_launch must be positive");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->tune.K = k;
        ctx->tune.np = pairs_per_lane;
        ctx->tune.schedule = schedule;
        ctx->tune.xcd_order = xcd_order ? 1 : 0;
        ctx->tune.min_waves = min_waves;
    });
}

int gsmc_context_release_buffers(gsmc_context* ctx)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        if (ctx->compute) HIP_OK(hipStreamSynchronize(ctx->compute));
        if (ctx->copy) HIP_OK(hipStreamSynchronize(ctx->copy));
        std::vector<gsmc_batch*> own{ctx->dev_batch.get()};
        for (auto& sl : ctx->slots) {
            HIP_OK(hipStreamSynchronize(sl->stream));
            own.push_back(sl->batch.get());
        }
        for (gsmc_batch* b : own) {
            if (!b) continue;
            if (b->pending) HIP_OK(hipEventSynchronize(b->done));
            b->pending = false;
            b->last.clear();  // the metadata goes with the buffers: the next launch uploads it again
            b->meta.release();
            b->ck.release();
            b->yb.release();
            b->part.release();
        }
        for (auto& sl : ctx->slots) sl->out.release();
        ctx->fwd_buf.release();
    });
}

int gsmc_context_set_state_budget(gsmc_context* ctx, uint64_t max_state_bytes)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->tune.max_state_bytes = max_state_bytes;
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
        check_flags(flags);
        require(ctx != nullptr, "context is NULL");
        require(P >= 0, "negative n_pairs");
        if (P == 0 || ctx->S == 0) return;
        require(mean != nullptr, "mean is NULL");
        require(!(flags & GSMC_CI) || (lower && upper), "lower/upper required with GSMC_CI");
        require(!(flags & GSMC_POSTERIOR) || (alpha && beta), "alpha/beta required with GSMC_POSTERIOR");
        check_pairs(ctx, pi, pj, P);
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        std::vector<Block> blocks{{0, (int32_t)ctx->S, 0, (int32_t)ctx->S}};
        float* outs[5];
        int k = 0;
        outs[k++] = mean;
        if (flags & GSMC_CI) outs[k++] = lower, outs[k++] = upper;
        if (flags & GSMC_POSTERIOR) outs[k++] = alpha, outs[k++] = beta;
        run_to_host(ctx, pi, pj, P, blocks, -1, flags, outs, 2);
    });
}

int gsmc_run_fb_blockwise(gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P, int64_t core,
                          int64_t flank, int64_t pair_batch, int max_streams, int flags, float* mean, float* lower,
                          float* upper, float* alpha, float* beta)
{
    return guarded([&] {
        check_flags(flags);
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
        run_to_host(ctx, pi, pj, P, blocks, pair_batch, flags, outs, max_streams);
    });
}

int gsmc_run_fb_device(gsmc_context* ctx, const int32_t* d_pi, const int32_t* d_pj, int64_t P, int64_t core,
                       int64_t flank, int flags, float* d_mean, float* d_lower, float* d_upper, float* d_alpha,
                       float* d_beta, void* stream)
{
    return guarded([&] {
        check_flags(flags);
        require(ctx != nullptr, "context is NULL");
        require(core > 0, "core_block_sites must be positive.");
        require(flank >= 0, "flank_sites must be non-negative.");
        require(P >= 0, "negative n_pairs");
        if (P == 0 || ctx->S == 0) return;
        require(d_mean != nullptr, "mean is NULL");
        require(!(flags & GSMC_CI) || (d_lower && d_upper), "lower/upper required with GSMC_CI");
        require(!(flags & GSMC_POSTERIOR) || (d_alpha && d_beta), "alpha/beta required with GSMC_POSTERIOR");
        require(!(flags & GSMC_SUMMARY), "use gsmc_run_fb_summary for per-site summaries");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        if (!ctx->dev_batch) ctx->dev_batch = new_batch(ctx);
        JobSpec js;
        js.blocks = make_blocks(ctx->S, core, flank);
        check_codes(flags, js.blocks, 0);
        js.pi = d_pi;
        js.pj = d_pj;
        js.P = P;
        js.out = Outputs{d_mean, d_lower, d_upper, d_alpha, d_beta};
        launch_plan(ctx->dev_batch.get(), {js}, flags, static_cast<hipStream_t>(stream));
    });
}

int gsmc_run_fb_device_noalloc(gsmc_context* ctx, const int32_t* d_pi, const int32_t* d_pj, int64_t P, int64_t core,
                               int64_t flank, int flags, float* d_mean, float* d_lower, float* d_upper, float* d_alpha,
                               float* d_beta, void* stream)
{
    return guarded([&] {
        check_flags(flags);
        require(ctx != nullptr, "context is NULL");
        require(core > 0, "core_block_sites must be positive.");
        require(flank >= 0, "flank_sites must be non-negative.");
        require(P >= 0, "negative n_pairs");
        if (P == 0 || ctx->S == 0) return;
        require(d_pi != nullptr && d_pj != nullptr && d_mean != nullptr, "NULL argument");
        require(!(flags & GSMC_CI) || (d_lower && d_upper), "lower/upper required with GSMC_CI");
        require(!(flags & GSMC_POSTERIOR) || (d_alpha && d_beta), "alpha/beta required with GSMC_POSTERIOR");
        require(!(flags & GSMC_SUMMARY), "use gsmc_run_fb_summary for per-site summaries");
        const int64_t S = ctx->S;
        const int64_t c = std::min(core, S);
        const int64_t nb = (S + c - 1) / c;
        require(!(flags & GSMC_Q4) || c % 2 == 0, "uint4 output needs an even core_block_sites");
        require(!((flags & (GSMC_Q8 | GSMC_Q4)) && (flags & (GSMC_CI | GSMC_POSTERIOR))),
                "quantized output cannot be combined with GSMC_CI or GSMC_POSTERIOR");
        require(!((flags & GSMC_Q8) && (flags & GSMC_Q4)), "GSMC_Q8 and GSMC_Q4 are exclusive");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        require(ctx->scratch != nullptr, "no device scratch for this context");
        // Checkpoints every 8 sites, one pair per lane, static grid: the
        // same kernels (and so the same results) as gsmc_run_fb_device.
        constexpr int K = 8;
        const int64_t tiles = (P + 63) / 64;
        const int64_t n_items = nb * tiles;
        require(nb <= INT32_MAX && n_items <= INT32_MAX, "too many work items");
        // scratch layout: job | blocks | items | flank states | checkpoints
        const int64_t rows = 2 * (((c + K - 1) / K + 1) / 2);  // checkpoint float2 rows per item
        const int64_t ck_stride = rows * 64;
        const size_t off_blocks = align256(sizeof(JobDev));
        const size_t off_items = off_blocks + align256((size_t)nb * sizeof(BlockDev));
        const size_t off_yb = off_items + align256((size_t)n_items * sizeof(ItemDev));
        const size_t off_ck = off_yb + align256((size_t)n_items * 64 * sizeof(float2));
        const size_t need = off_ck + (size_t)n_items * (size_t)ck_stride * sizeof(float2);
        if (need > kScratchBytes)
            throw Error(GSMC_EINVAL, "gsmc_run_fb_device_noalloc: this call needs " + std::to_string(need) +
                                         " bytes of device scratch, above its " + std::to_string(kScratchBytes) +
                                         " byte maximum (fewer pairs or sites per call, or use gsmc_run_fb_device)");
        // first site's gap from position -1 (full stretches from the prior
        // table); block restarts step one bp from (position - 1)
        const uint32_t full = descriptor(round_steps(ctx->pos[0] + 1.0), ctx->N);
        const int64_t q = full >> kDescLayerBits;
        const uint32_t sd0 = q < ctx->prior_n ? (full & kDescLayerMask) : full;
        const int32_t q0 = q < ctx->prior_n ? (int32_t)q : 0;
        uint32_t sd_restart = descriptor(1, ctx->N);
        for (int64_t b = 0; b < nb; ++b) {
            const int64_t ps = std::max<int64_t>(0, b * c - flank);
            if (ps == 0) continue;
            const double pp = ctx->pos[(size_t)ps];
            require(descriptor(round_steps(pp - (pp - 1.0)), ctx->N) == sd_restart,
                    "gsmc_run_fb_device_noalloc: positions too large for one-bp block restarts");
        }
        char* base = static_cast<char*>(ctx->scratch);
        JobDev job{d_mean, d_lower, d_upper, d_alpha, d_beta, d_pi, d_pj, P, 0, S, -1};
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int64_t threads = std::max<int64_t>(n_items, nb);
        hipLaunchKernelGGL(build_meta_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, job,
                           reinterpret_cast<JobDev*>(base), reinterpret_cast<BlockDev*>(base + off_blocks),
                           reinterpret_cast<ItemDev*>(base + off_items), S, c, flank, (int32_t)nb, (int32_t)tiles, sd0,
                           q0, sd_restart, (int32_t)K, ck_stride);
        HIP_OK(hipGetLastError());
        FbArgs a{};
        a.packed = ctx->packed.as<uint64_t>();
        a.n_haps = ctx->n;
        a.W = ctx->W;
        a.desc = ctx->desc.as<uint32_t>();
        a.prior = ctx->prior.as<float2>();
        a.qbits = (flags & GSMC_Q8) ? 8 : (flags & GSMC_Q4) ? 4 : 0;
        a.q_inv_range = 1.0f / logf(20.0f * ctx->Ne);
        a.jobs = reinterpret_cast<const JobDev*>(base);
        a.blocks = reinterpret_cast<const BlockDev*>(base + off_blocks);
        a.items = reinterpret_cast<const ItemDev*>(base + off_items);
        a.n_items = n_items;
        a.yb = reinterpret_cast<float2*>(base + off_yb);
        a.ck = reinterpret_cast<float2*>(base + off_ck);
        a.t = ctx->tabs->view;
        a.unscale = 2.0f * ctx->Ne;
        Tuning tu;
        tu.K = K;
        tu.np = 1;
        tu.schedule = 0;
        const KernelPair kp = kernels_for(tu, flags);
        launch(kp.fwd_items, (unsigned)n_items, a, st);
        launch(kp.bwd_items, (unsigned)n_items, a, st);
    });
}

int gsmc_run_fb_summary(gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P, int64_t core,
                        int64_t flank, float* site_mean, float* site_min, float* site_max)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        require(P >= 0, "negative n_pairs");
        require(flank >= 0, "flank_sites must be non-negative.");
        if (P == 0 || ctx->S == 0) return;
        require(pi != nullptr && pj != nullptr && site_mean != nullptr, "NULL argument");
        check_pairs(ctx, pi, pj, P);
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        if (!ctx->dev_batch) ctx->dev_batch = new_batch(ctx);
        const int64_t S = ctx->S;
        const int64_t c = core > 0 ? core : S;
        const int64_t fl = core > 0 ? flank : 0;
        // Pairs per chunk: partials (16 B per 64-pair group and site) and
        // checkpoints (8 B per pair every K sites) within half of the memory
        // the launch can use: free memory plus the workspace buffers it will
        // reuse (a workspace keeps its largest earlier launch's buffers).
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b));
        const double avail = (double)free_b + (double)ctx->dev_batch->ck.bytes + (double)ctx->dev_batch->part.bytes;
        const double per_pair = 16.0 / 64.0 * (double)S + 8.0 * (double)(S + 2 * fl) / ctx->tune.K + 64.0;
        int64_t chunk = (int64_t)(0.5 * avail / per_pair) / 128 * 128;
        chunk = std::max<int64_t>(chunk, 128);
        chunk = std::min<int64_t>(chunk, (P + 127) / 128 * 128);
        ctx->sum_pairs.ensure((size_t)chunk * 2 * sizeof(int32_t), ctx->device);
        ctx->sum_site.ensure((size_t)S * (3 * sizeof(float) + sizeof(double) + 2 * sizeof(float)), ctx->device);
        int32_t* d_pi = ctx->sum_pairs.as<int32_t>();
        int32_t* d_pj = d_pi + chunk;
        double* tot = ctx->sum_site.as<double>();
        float* d_mean = reinterpret_cast<float*>(tot + S);
        float* d_min = d_mean + S;
        float* d_max = d_min + S;
        float* tmin = d_max + S;
        float* tmax = tmin + S;
        for (int64_t p0 = 0; p0 < P; p0 += chunk) {
            const int64_t n = std::min(chunk, P - p0);
            // the previous chunk's kernels read the pair buffer
            HIP_OK(hipStreamSynchronize(ctx->compute));
            HIP_OK(hipMemcpyAsync(d_pi, pi + p0, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->compute));
            HIP_OK(hipMemcpyAsync(d_pj, pj + p0, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->compute));
            JobSpec js;
            js.blocks = make_blocks(S, c, fl);
            js.pi = d_pi;
            js.pj = d_pj;
            js.P = n;
            js.out.site_mean = d_mean;
            js.out.site_min = d_min;
            js.out.site_max = d_max;
            js.sum_first = p0 == 0;
            js.sum_last = p0 + n == P;
            js.tot = tot;
            js.tmin = tmin;
            js.tmax = tmax;
            js.P_total = P;
            launch_plan(ctx->dev_batch.get(), {js}, GSMC_SUMMARY, ctx->compute);
        }
        HIP_OK(hipMemcpyAsync(site_mean, d_mean, (size_t)S * sizeof(float), hipMemcpyDeviceToHost, ctx->compute));
        if (site_min)
            HIP_OK(hipMemcpyAsync(site_min, d_min, (size_t)S * sizeof(float), hipMemcpyDeviceToHost, ctx->compute));
        if (site_max)
            HIP_OK(hipMemcpyAsync(site_max, d_max, (size_t)S * sizeof(float), hipMemcpyDeviceToHost, ctx->compute));
        HIP_OK(hipStreamSynchronize(ctx->compute));
    });
}

namespace gsmc {
namespace {

// Work layout of a per-gene statistics call: the genes' boundaries and the
// block cores' boundaries cut the sites into pieces; the pieces some gene
// covers are numbered in site order, and site_piece gives each site a code:
// 2 k + 1 in piece k, 2 k in no gene after k pieces (non-decreasing along
// the sites). A gene covers a contiguous run of pieces, and each piece lies
// inside one block core, so one work item per 64 pairs writes its sums.
struct GenePlan {
    std::vector<int32_t> site_piece;
    int64_t n_pieces = 0;
    std::vector<int32_t> gene_pb, gene_pe;
    std::vector<int64_t> gene_n;
};

static void plan_genes(int64_t S, int64_t core, int64_t n_genes, const int64_t* lo, const int64_t* hi,
                       GenePlan& gp)
{
    std::vector<std::pair<int64_t, int>> ev;  // (site, +1 gene starts / -1 gene ends)
    ev.reserve((size_t)n_genes * 2);
    for (int64_t g = 0; g < n_genes; ++g) {
        if (!(lo[g] >= 0 && lo[g] <= hi[g] && hi[g] <= S))
            throw Error(GSMC_EINVAL, "gene " + std::to_string(g) + " site range [" + std::to_string(lo[g]) + ", " +
                                         std::to_string(hi[g]) + ") is not within [0, " + std::to_string(S) + "]");
        if (hi[g] > lo[g]) {
            ev.emplace_back(lo[g], 1);
            ev.emplace_back(hi[g], -1);
        }
    }
    std::sort(ev.begin(), ev.end());
    gp.site_piece.assign((size_t)S, 0);
    std::vector<int64_t> piece_lo;
    int64_t done = 0;  // sites [0, done) have their codes
    int cover = 0;
    for (size_t i = 0; i < ev.size();) {
        const int64_t at = ev[i].first;
        while (i < ev.size() && ev[i].first == at) cover += ev[i++].second;
        if (cover > 0 && i < ev.size()) {
            const int64_t end = ev[i].first;  // next boundary: [at, end) is covered
            for (int64_t s = at; s < end;) {
                const int64_t e = std::min(end, (s / core + 1) * core);  // up to the block core's end
                require(piece_lo.size() < (size_t)(INT32_MAX / 2), "too many gene pieces");
                const int32_t k = (int32_t)piece_lo.size();
                std::fill(gp.site_piece.begin() + done, gp.site_piece.begin() + s, 2 * k);
                std::fill(gp.site_piece.begin() + s, gp.site_piece.begin() + e, 2 * k + 1);
                done = e;
                piece_lo.push_back(s);
                s = e;
            }
        }
    }
    std::fill(gp.site_piece.begin() + done, gp.site_piece.end(), 2 * (int32_t)piece_lo.size());
    gp.n_pieces = (int64_t)piece_lo.size();
    gp.gene_pb.resize((size_t)n_genes);
    gp.gene_pe.resize((size_t)n_genes);
    gp.gene_n.resize((size_t)n_genes);
    for (int64_t g = 0; g < n_genes; ++g) {
        gp.gene_pb[(size_t)g] = (int32_t)(std::lower_bound(piece_lo.begin(), piece_lo.end(), lo[g]) - piece_lo.begin());
        gp.gene_pe[(size_t)g] = (int32_t)(std::lower_bound(piece_lo.begin(), piece_lo.end(), hi[g]) - piece_lo.begin());
        gp.gene_n[(size_t)g] = hi[g] - lo[g];
    }
}
}  // namespace
}  // namespace gsmc

int gsmc_run_fb_gene_stats(gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P, int64_t core,
                           int64_t flank, int64_t pair_batch, int64_t n_genes, const int64_t* gene_lo,
                           const int64_t* gene_hi, int n_bins,
                           const double* edges, int64_t* count, double* lin_sum, double* log_sum, double* log_sq_sum,
                           double* min_lin, double* min_log, int64_t* histogram)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        require(P >= 0, "negative n_pairs");
        require(n_genes >= 0, "negative n_genes");
        require(flank >= 0, "flank_sites must be non-negative.");
        require(n_bins >= 1 && n_bins <= kMaxGeneBins,
                "n_bins must be in [1, " + std::to_string(kMaxGeneBins) + "]");
        require(edges != nullptr, "edges is NULL");
        for (int i = 0; i <= n_bins; ++i) require(std::isfinite(edges[i]), "histogram edges must be finite");
        for (int i = 0; i < n_bins; ++i) require(edges[i] < edges[i + 1], "histogram edges must increase");
        if (n_genes == 0) return;
        require(gene_lo != nullptr && gene_hi != nullptr && count != nullptr && lin_sum != nullptr &&
                    log_sum != nullptr && log_sq_sum != nullptr && min_lin != nullptr && min_log != nullptr &&
                    histogram != nullptr,
                "NULL argument");
        require(n_genes <= INT32_MAX, "too many genes");
        const int64_t S = ctx->S;
        const int64_t c = core > 0 ? std::min(core, std::max<int64_t>(S, 1)) : std::max<int64_t>(S, 1);
        const int64_t fl = core > 0 ? flank : 0;
        GenePlan gp;
        plan_genes(S, c, n_genes, gene_lo, gene_hi, gp);
        if (P == 0 || S == 0 || gp.n_pieces == 0) return;
        require(pi != nullptr && pj != nullptr, "NULL argument");
        check_pairs(ctx, pi, pj, P);
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        if (!ctx->dev_batch) ctx->dev_batch = new_batch(ctx);
        const int64_t n_pieces = gp.n_pieces;
        // Pairs per chunk: piece sums (16 B per piece) and checkpoints (8 B
        // every K sites of each padded block) within 70 % of the memory the
        // call can use (free memory plus the buffers it reuses); chunks of
        // equal size, so that no launch is a small remainder that leaves
        // most of the chip idle.
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b));
        const double avail = (double)free_b + (double)ctx->dev_batch->ck.bytes + (double)ctx->gene_part.bytes;
        const int64_t nb = (S + c - 1) / c;
        const double per_pair = 16.0 * (double)n_pieces + 8.0 * (double)(S + 2 * fl * nb) / window_k(ctx->tune) + 64.0;
        int64_t most = std::max<int64_t>(64, (int64_t)(0.7 * avail / per_pair) / 64 * 64);
        if (pair_batch > 0) most = std::min(most, pair_batch);
        const int64_t n_chunks = (P + most - 1) / most;
        const int64_t per = (P + n_chunks - 1) / n_chunks;
        const int64_t chunk = pair_batch > 0 ? per : (per + 63) / 64 * 64;
        // device metadata: site pieces, genes, edges; accumulators
        const size_t o_pb = align256((size_t)S * 4);
        const size_t o_pe = o_pb + align256((size_t)n_genes * 4);
        const size_t o_n = o_pe + align256((size_t)n_genes * 4);
        const size_t o_edge = o_n + align256((size_t)n_genes * 8);
        const size_t meta_bytes = o_edge + (size_t)(n_bins + 1) * 8;
        const size_t a_lin = align256((size_t)n_genes * 8);
        const size_t a_log = a_lin + align256((size_t)n_genes * 8);
        const size_t a_sq = a_log + align256((size_t)n_genes * 8);
        const size_t a_mlin = a_sq + align256((size_t)n_genes * 8);
        const size_t a_mlog = a_mlin + align256((size_t)n_genes * 8);
        const size_t a_hist = a_mlog + align256((size_t)n_genes * 8);
        const size_t acc_bytes = a_hist + (size_t)n_genes * (size_t)n_bins * 8;
        // the previous call's kernels may still read these buffers
        HIP_OK(hipStreamSynchronize(ctx->compute));
        ctx->gene_meta.ensure(meta_bytes, ctx->device, true);
        ctx->gene_acc.ensure(acc_bytes, ctx->device, true);
        ctx->gene_pairs.ensure((size_t)chunk * 2 * sizeof(int32_t), ctx->device, true);
        ctx->gene_part.ensure((size_t)chunk * (size_t)n_pieces * sizeof(double2), ctx->device, true);
        // workgroups per gene in the reduction: about 2048 in all, each over
        // at least 2048 pairs of the largest launch
        const int32_t n_split = (int32_t)std::max<int64_t>(
            1, std::min<int64_t>((2048 + n_genes - 1) / n_genes, std::max<int64_t>(1, chunk / 2048)));
        require(n_split <= 65535, "internal: gene reduction split");
        ctx->gene_split.ensure((size_t)n_genes * (size_t)n_split * (5 * sizeof(double) + (size_t)n_bins * 4),
                               ctx->device, true);
        char* meta = ctx->gene_meta.as<char>();
        char* acc = ctx->gene_acc.as<char>();
        auto h2d = [&](void* dst, const void* src, size_t bytes) {
            if (bytes) HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->compute));
        };
        h2d(meta, gp.site_piece.data(), (size_t)S * 4);
        h2d(meta + o_pb, gp.gene_pb.data(), (size_t)n_genes * 4);
        h2d(meta + o_pe, gp.gene_pe.data(), (size_t)n_genes * 4);
        h2d(meta + o_n, gp.gene_n.data(), (size_t)n_genes * 8);
        h2d(meta + o_edge, edges, (size_t)(n_bins + 1) * 8);
        h2d(acc, count, (size_t)n_genes * 8);
        h2d(acc + a_lin, lin_sum, (size_t)n_genes * 8);
        h2d(acc + a_log, log_sum, (size_t)n_genes * 8);
        h2d(acc + a_sq, log_sq_sum, (size_t)n_genes * 8);
        h2d(acc + a_mlin, min_lin, (size_t)n_genes * 8);
        h2d(acc + a_mlog, min_log, (size_t)n_genes * 8);
        h2d(acc + a_hist, histogram, (size_t)n_genes * (size_t)n_bins * 8);

        GeneReduceArgs gr;
        gr.part = ctx->gene_part.as<double2>();
        gr.gene_pb = reinterpret_cast<const int32_t*>(meta + o_pb);
        gr.gene_pe = reinterpret_cast<const int32_t*>(meta + o_pe);
        gr.gene_n = reinterpret_cast<const int64_t*>(meta + o_n);
        gr.edges = reinterpret_cast<const double*>(meta + o_edge);
        gr.n_bins = n_bins;
        gr.count = reinterpret_cast<int64_t*>(acc);
        gr.lin_sum = reinterpret_cast<double*>(acc + a_lin);
        gr.log_sum = reinterpret_cast<double*>(acc + a_log);
        gr.log_sq_sum = reinterpret_cast<double*>(acc + a_sq);
        gr.min_lin = reinterpret_cast<double*>(acc + a_mlin);
        gr.min_log = reinterpret_cast<double*>(acc + a_mlog);
        gr.hist = reinterpret_cast<int64_t*>(acc + a_hist);
        gr.n_split = n_split;
        gr.split_val = ctx->gene_split.as<double>();
        gr.split_hist = reinterpret_cast<uint32_t*>(gr.split_val + (size_t)n_genes * (size_t)n_split * 5);

        int32_t* d_pi = ctx->gene_pairs.as<int32_t>();
        int32_t* d_pj = d_pi + chunk;
        for (int64_t p0 = 0; p0 < P; p0 += chunk) {
            const int64_t n = std::min(chunk, P - p0);
            // the previous chunk's kernels read the pair buffer
            HIP_OK(hipStreamSynchronize(ctx->compute));
            h2d(d_pi, pi + p0, (size_t)n * sizeof(int32_t));
            h2d(d_pj, pj + p0, (size_t)n * sizeof(int32_t));
            JobSpec js;
            js.blocks = make_blocks(S, c, fl);
            js.pi = d_pi;
            js.pj = d_pj;
            js.P = n;
            js.gene_piece = reinterpret_cast<const int32_t*>(meta);
            js.gene_part = ctx->gene_part.as<double2>();
            launch_plan(ctx->dev_batch.get(), {js}, kGeneLaunch, ctx->compute);
            gr.P = n;
            hipLaunchKernelGGL(gene_reduce_kernel, dim3((unsigned)n_genes, (unsigned)n_split), dim3(256), 0,
                               ctx->compute, gr);
            hipLaunchKernelGGL(gene_combine_kernel, dim3((unsigned)n_genes), dim3(256), 0, ctx->compute, gr);
            HIP_OK(hipGetLastError());
        }
        auto d2h = [&](void* dst, const void* src, size_t bytes) {
            HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->compute));
        };
        d2h(count, acc, (size_t)n_genes * 8);
        d2h(lin_sum, acc + a_lin, (size_t)n_genes * 8);
        d2h(log_sum, acc + a_log, (size_t)n_genes * 8);
        d2h(log_sq_sum, acc + a_sq, (size_t)n_genes * 8);
        d2h(min_lin, acc + a_mlin, (size_t)n_genes * 8);
        d2h(min_log, acc + a_mlog, (size_t)n_genes * 8);
        d2h(histogram, acc + a_hist, (size_t)n_genes * (size_t)n_bins * 8);
        HIP_OK(hipStreamSynchronize(ctx->compute));
    });
}

namespace gsmc {
namespace {
static void launch_fwd_only(gsmc_context* ctx, const int32_t* d_pi, const int32_t* d_pj, int64_t P, int flags,
                     float* d_mean, float* d_lower, float* d_upper, hipStream_t stream)
{
    FwdArgs a;
    a.packed = ctx->packed.as<uint64_t>();
    a.n_haps = ctx->n;
    a.W = ctx->W;
    a.S = ctx->S;
    a.desc = ctx->desc_h2.as<uint32_t>();
    a.h2q = ctx->tabs->h2q;
    a.N = ctx->N;
    a.pair_i = d_pi;
    a.pair_j = d_pj;
    a.P = P;
    a.two_ne = 2.0f * ctx->Ne;
    a.mean = d_mean;
    a.lower = d_lower;
    a.upper = d_upper;
    const dim3 grid((unsigned)((P + 63) / 64));
    if (flags & GSMC_CI)
        hipLaunchKernelGGL(fwd_only_kernel<true>, grid, dim3(64), 0, stream, a);
    else
        hipLaunchKernelGGL(fwd_only_kernel<false>, grid, dim3(64), 0, stream, a);
    HIP_OK(hipGetLastError());
}
}  // namespace
}  // namespace gsmc

int gsmc_run_fwd_device(gsmc_context* ctx, const int32_t* d_pi, const int32_t* d_pj, int64_t P, int flags,
                        float* d_mean, float* d_lower, float* d_upper, void* stream)
{
    return guarded([&] {
        check_flags(flags);
        require(ctx != nullptr, "context is NULL");
        require(P >= 0, "negative n_pairs");
        if (P == 0 || ctx->S == 0) return;
        require(d_pi != nullptr && d_pj != nullptr && d_mean != nullptr, "NULL argument");
        require(!(flags & GSMC_CI) || (d_lower && d_upper), "lower/upper required with GSMC_CI");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        launch_fwd_only(ctx, d_pi, d_pj, P, flags, d_mean, d_lower, d_upper, static_cast<hipStream_t>(stream));
    });
}

int gsmc_run_fwd(gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P, int flags, float* mean,
                 float* lower, float* upper)
{
    return guarded([&] {
        check_flags(flags);
        require(ctx != nullptr, "context is NULL");
        require(P >= 0, "negative n_pairs");
        if (P == 0 || ctx->S == 0) return;
        require(pi != nullptr && pj != nullptr && mean != nullptr, "NULL argument");
        const bool ci = (flags & GSMC_CI) != 0;
        require(!ci || (lower && upper), "lower/upper required with GSMC_CI");
        check_pairs(ctx, pi, pj, P);
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        const int64_t S = ctx->S;
        const int nout = ci ? 3 : 1;
        // Pairs per chunk: outputs of a chunk within a quarter of free memory.
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b));
        const double per_pair = 4.0 * nout * (double)S + 8.0;
        int64_t chunk = (int64_t)(0.25 * (double)free_b / per_pair) / 64 * 64;
        chunk = std::min<int64_t>(std::max<int64_t>(chunk, 64), (P + 63) / 64 * 64);
        ctx->fwd_buf.ensure((size_t)chunk * 8 + (size_t)nout * (size_t)S * (size_t)chunk * 4, ctx->device);
        int32_t* d_pi = ctx->fwd_buf.as<int32_t>();
        int32_t* d_pj = d_pi + chunk;
        float* d_out = reinterpret_cast<float*>(d_pj + chunk);
        float* host[3] = {mean, lower, upper};
        for (int64_t p0 = 0; p0 < P; p0 += chunk) {
            const int64_t n = std::min(chunk, P - p0);
            HIP_OK(hipStreamSynchronize(ctx->compute));
            HIP_OK(hipMemcpyAsync(d_pi, pi + p0, (size_t)n * 4, hipMemcpyHostToDevice, ctx->compute));
            HIP_OK(hipMemcpyAsync(d_pj, pj + p0, (size_t)n * 4, hipMemcpyHostToDevice, ctx->compute));
            float* dev[3] = {d_out, d_out + (size_t)S * n, d_out + 2 * (size_t)S * n};
            launch_fwd_only(ctx, d_pi, d_pj, n, flags, dev[0], ci ? dev[1] : nullptr, ci ? dev[2] : nullptr,
                            ctx->compute);
            for (int o = 0; o < nout; ++o)
                HIP_OK(hipMemcpy2DAsync(host[o] + p0, (size_t)P * 4, dev[o], (size_t)n * 4, (size_t)n * 4, (size_t)S,
                                        hipMemcpyDeviceToHost, ctx->compute));
        }
        HIP_OK(hipStreamSynchronize(ctx->compute));
    });
}

int gsmc_run_fwd_states(gsmc_context* ctx, const int32_t* pi, const int32_t* pj, int64_t P, float* mean_log10,
                        float* cv_log10)
{
    return guarded([&] {
        require(ctx != nullptr, "context is NULL");
        require(P >= 0, "negative n_pairs");
        if (P == 0 || ctx->S == 0) return;
        require(pi != nullptr && pj != nullptr && mean_log10 != nullptr && cv_log10 != nullptr, "NULL argument");
        check_pairs(ctx, pi, pj, P);
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        init_streams(ctx);
        const int64_t S = ctx->S;
        // first site: its gap from position -1, full stretches from the prior table
        const uint32_t full = descriptor(round_steps(ctx->pos[0] + 1.0), ctx->N);
        const int64_t q = full >> kDescLayerBits;
        const int32_t q0 = q < ctx->prior_n ? (int32_t)q : 0;
        const uint32_t sd0 = q < ctx->prior_n ? (full & kDescLayerMask) : full;
        // pairs per chunk: the two [S][chunk] planes within a quarter of free memory
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b));
        int64_t chunk = (int64_t)(0.25 * (double)free_b / (8.0 * (double)S + 8.0)) / 64 * 64;
        chunk = std::min<int64_t>(std::max<int64_t>(chunk, 64), (P + 63) / 64 * 64);
        ctx->fwd_buf.ensure((size_t)chunk * 8 + 2 * (size_t)S * (size_t)chunk * 4, ctx->device);
        int32_t* d_pi = ctx->fwd_buf.as<int32_t>();
        int32_t* d_pj = d_pi + chunk;
        float* d_m = reinterpret_cast<float*>(d_pj + chunk);
        FbArgs a{};
        a.packed = ctx->packed.as<uint64_t>();
        a.n_haps = ctx->n;
        a.W = ctx->W;
        a.desc = ctx->desc.as<uint32_t>();
        a.prior = ctx->prior.as<float2>();
        a.t = ctx->tabs->view;
        for (int64_t p0 = 0; p0 < P; p0 += chunk) {
            const int64_t n = std::min(chunk, P - p0);
            float* d_c = d_m + (size_t)S * n;
            HIP_OK(hipStreamSynchronize(ctx->compute));
            HIP_OK(hipMemcpyAsync(d_pi, pi + p0, (size_t)n * 4, hipMemcpyHostToDevice, ctx->compute));
            HIP_OK(hipMemcpyAsync(d_pj, pj + p0, (size_t)n * 4, hipMemcpyHostToDevice, ctx->compute));
            hipLaunchKernelGGL(fwd_states_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, ctx->compute, a,
                               d_pi, d_pj, n, S, q0, sd0, d_m, d_c);
            HIP_OK(hipGetLastError());
            HIP_OK(hipMemcpy2DAsync(mean_log10 + p0, (size_t)P * 4, d_m, (size_t)n * 4, (size_t)n * 4, (size_t)S,
                                    hipMemcpyDeviceToHost, ctx->compute));
            HIP_OK(hipMemcpy2DAsync(cv_log10 + p0, (size_t)P * 4, d_c, (size_t)n * 4, (size_t)n * 4, (size_t)S,
                                    hipMemcpyDeviceToHost, ctx->compute));
        }
        HIP_OK(hipStreamSynchronize(ctx->compute));
    });
}

int gsmc_lookup_probe(gsmc_context* ctx, int64_t waves, int32_t steps, void* out_scratch, void* stream, float* ms)
{
    return guarded([&] {
        require(ctx != nullptr && out_scratch != nullptr && ms != nullptr, "NULL argument");
        require(waves > 0 && waves <= INT32_MAX / 64 && steps > 0, "bad probe shape");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        hipStream_t st = static_cast<hipStream_t>(stream);
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        float* out = static_cast<float*>(out_scratch);
        // one untimed launch (page-in), then the timed one
        hipLaunchKernelGGL(lookup_probe_kernel, dim3((unsigned)waves), dim3(64), 0, st, ctx->tabs->view, ctx->N,
                           std::min<int32_t>(steps, 64), out);
        HIP_OK(hipEventRecord(e0, st));
        hipLaunchKernelGGL(lookup_probe_kernel, dim3((unsigned)waves), dim3(64), 0, st, ctx->tabs->view, ctx->N, steps,
                           out);
        HIP_OK(hipEventRecord(e1, st));
        const hipError_t le = hipGetLastError();
        HIP_OK(hipEventSynchronize(e1));
        HIP_OK(hipEventElapsedTime(ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIP_OK(le);
    });
}

int gsmc_get_device(int* device)
{
    return guarded([&] {
        require(device != nullptr, "NULL argument");
        HIP_OK(hipGetDevice(device));
    });
}

int gsmc_batch_create(gsmc_context* ctx, gsmc_batch** out)
{
    return guarded([&] {
        require(ctx != nullptr && out != nullptr, "NULL argument");
        HIP_OK(hipSetDevice(ctx->device));
        *out = new_batch(ctx).release();
    });
}

void gsmc_batch_destroy(gsmc_batch* b) { delete b; }

int gsmc_batch_reserve(gsmc_batch* b, uint64_t checkpoint_bytes, int64_t max_items, int64_t max_jobs)
{
    return guarded([&] {
        require(b != nullptr, "batch is NULL");
        require(max_items >= 0 && max_jobs >= 0, "negative reservation");
        gsmc_context* ctx = b->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        if (b->pending) HIP_OK(hipEventSynchronize(b->done));
        // A launch's metadata: jobs, then at most one block per item, then items.
        const size_t meta = align256((size_t)max_jobs * sizeof(JobDev)) +
                            align256((size_t)max_items * sizeof(BlockDev)) + (size_t)max_items * sizeof(ItemDev);
        b->pending = false;
        b->meta.ensure(meta, ctx->device, true);
        b->stage.ensure(meta);
        b->ck.ensure(std::max<uint64_t>(checkpoint_bytes, 1), ctx->device, true);
        b->queue.ensure(256, ctx->device, true);
        b->yb.ensure(yb_bytes(max_items, kMaxPairsPerLane), ctx->device, true);
        b->reserved = true;
    });
}

// C++ linkage for helpers that return C++ types, inside the C ABI block
extern "C++" {
namespace gsmc {
namespace {
std::vector<JobSpec> specs_of(const gsmc_context* ctx, const gsmc_job* jobs, int64_t n_jobs, int flags);
}
}  // namespace gsmc
}  // extern "C++"

// A captured launch: private buffers with the plan's metadata uploaded, and
// an executable graph of the launch's kernels.
struct gsmc_graph {
    gsmc_context* ctx = nullptr;
    std::unique_ptr<gsmc_batch> buf;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~gsmc_graph()
    {
        if (ctx) (void)hipSetDevice(ctx->device);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

int gsmc_batch_capture(gsmc_context* ctx, const gsmc_job* jobs, int64_t n_jobs, int flags, gsmc_graph** out)
{
    return guarded([&] {
        check_flags(flags);
        require(ctx != nullptr && out != nullptr, "NULL argument");
        require(n_jobs >= 0 && (jobs != nullptr || n_jobs == 0), "bad job list");
        require(!((flags & GSMC_SUMMARY) && (flags & (GSMC_CI | GSMC_POSTERIOR))),
                "GSMC_SUMMARY cannot be combined with GSMC_CI or GSMC_POSTERIOR");
        std::vector<JobSpec> specs = specs_of(ctx, jobs, n_jobs, flags);
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        auto g = std::make_unique<gsmc_graph>();
        g->ctx = ctx;
        g->buf = new_batch(ctx);
        const bool summary = (flags & GSMC_SUMMARY) != 0;
        const Tuning tu = launch_tuning(ctx->tune, flags);
        Plan pl = make_plan(ctx, specs, tu, launch_np(tu, flags), summary, stored_states(tu, flags), 0, flags);
        require(pl.n_items > 0, "nothing to capture: every job is empty");
        ensure_buffers(g->buf.get(), pl, summary);
        HIP_OK(hipMemcpy(g->buf->meta.p, pl.bytes.data(), pl.bytes.size(), hipMemcpyHostToDevice));
        const FbArgs a = make_args(g->buf.get(), pl, flags);
        hipStream_t s = nullptr;
        HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_kernels(ctx, pl, a, flags, s);
        } catch (...) {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(s, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            (void)hipStreamDestroy(s);
            throw;
        }
        const hipError_t e = hipStreamEndCapture(s, &g->graph);
        (void)hipStreamDestroy(s);
        HIP_OK(e);
        HIP_OK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        *out = g.release();
    });
}

int gsmc_graph_launch(gsmc_graph* g, void* stream)
{
    return guarded([&] {
        require(g != nullptr, "graph is NULL");
        HIP_OK(hipSetDevice(g->ctx->device));
        HIP_OK(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
    });
}

void gsmc_graph_destroy(gsmc_graph* g) { delete g; }

extern "C++" {
namespace gsmc {
namespace {
std::vector<JobSpec> specs_of(const gsmc_context* ctx, const gsmc_job* jobs, int64_t n_jobs, int flags)
{
    std::vector<JobSpec> specs;
    specs.reserve((size_t)n_jobs);
    for (int64_t i = 0; i < n_jobs; ++i) {
        const gsmc_job& j = jobs[i];
        require(j.site_lo >= 0 && j.site_lo <= j.site_hi && j.site_hi <= ctx->S, "job site range out of bounds");
        require(j.n_pairs >= 0, "negative n_pairs");
        require(j.flank_sites >= 0, "flank_sites must be non-negative.");
        if (j.n_pairs == 0 || j.site_hi == j.site_lo) continue;
        require(j.pair_i != nullptr && j.pair_j != nullptr, "job pointers are NULL");
        require((flags & GSMC_SUMMARY) ? j.site_mean != nullptr : j.mean != nullptr, "job outputs are NULL");
        require(!(flags & GSMC_CI) || (j.lower && j.upper), "lower/upper required with GSMC_CI");
        require(!(flags & GSMC_POSTERIOR) || (j.alpha && j.beta), "alpha/beta required with GSMC_POSTERIOR");
        const int64_t len = j.site_hi - j.site_lo;
        JobSpec js;
        js.seq_lo = j.site_lo;
        js.blocks = make_blocks(len, j.core_block_sites > 0 ? j.core_block_sites : len,
                                j.core_block_sites > 0 ? j.flank_sites : 0);
        for (Block& bk : js.blocks) {
            bk.cs += (int32_t)j.site_lo;
            bk.ce += (int32_t)j.site_lo;
            bk.ps += (int32_t)j.site_lo;
            bk.pe += (int32_t)j.site_lo;
        }
        js.pi = j.pair_i;
        js.pj = j.pair_j;
        js.P = j.n_pairs;
        js.out = Outputs{j.mean, j.lower, j.upper, j.alpha, j.beta, j.site_mean, j.site_min, j.site_max};
        js.row0 = j.site_lo;
        check_codes(flags, js.blocks, js.row0);
        specs.push_back(std::move(js));
    }
    return specs;
}
}  // namespace
}  // namespace gsmc
}  // extern "C++"

int gsmc_batch_launch(gsmc_batch* b, const gsmc_job* jobs, int64_t n_jobs, int flags, void* stream)
{
    return guarded([&] {
        check_flags(flags);
        require(b != nullptr, "batch is NULL");
        require(n_jobs >= 0 && (jobs != nullptr || n_jobs == 0), "bad job list");
        gsmc_context* ctx = b->ctx;
        require(!((flags & GSMC_SUMMARY) && (flags & (GSMC_CI | GSMC_POSTERIOR))),
                "GSMC_SUMMARY cannot be combined with GSMC_CI or GSMC_POSTERIOR");
        std::vector<JobSpec> specs = specs_of(ctx, jobs, n_jobs, flags);
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        launch_plan(b, specs, flags, static_cast<hipStream_t>(stream));
    });
}

int gsmc_batch_buffer_bytes(gsmc_batch* b, uint64_t* checkpoint_bytes)
{
    return guarded([&] {
        require(b != nullptr && checkpoint_bytes != nullptr, "NULL argument");
        *checkpoint_bytes = b->ck.bytes;
    });
}

int gsmc_batch_wait(gsmc_batch* b)
{
    return guarded([&] {
        require(b != nullptr, "batch is NULL");
        if (b->pending) HIP_OK(hipEventSynchronize(b->done));
        b->pending = false;
    });
}

int gsmc_batch_query(gsmc_batch* b, int* done)
{
    return guarded([&] {
        require(b != nullptr && done != nullptr, "NULL argument");
        if (!b->pending) {
            *done = 1;
            return;
        }
        const hipError_t e = hipEventQuery(b->done);
        if (e == hipErrorNotReady) {
            *done = 0;
            return;
        }
        HIP_OK(e);
        *done = 1;
    });
}

int gsmc_checkpoint_bytes(gsmc_context* ctx, int64_t n_sites, int64_t core, int64_t flank, int64_t n_pairs,
                          uint64_t* bytes)
{
    return guarded([&] {
        require(ctx != nullptr && bytes != nullptr, "NULL argument");
        const int64_t tile_pairs = 64 * (int64_t)launch_np(ctx->tune, 0);
        const int64_t tiles = (n_pairs + tile_pairs - 1) / tile_pairs;
        uint64_t windows = 0;  // state rows: one per checkpoint window (an even count per block), or per core site with K = 1
        for (const Block& b : make_blocks(n_sites, core > 0 ? core : std::max<int64_t>(n_sites, 1), core > 0 ? flank : 0))
            windows += ctx->tune.K == 1 ? (uint64_t)(2 * ((b.ce - b.cs + 1) / 2))
                                        : (uint64_t)(2 * (((b.ce - b.cs + ctx->tune.K - 1) / ctx->tune.K + 1) / 2));
        *bytes = windows * (uint64_t)tiles * (uint64_t)tile_pairs * sizeof(float2);
    });
}

int gsmc_device_copy(void* dst, const void* src, uint64_t bytes, void* stream)
{
    return guarded([&] {
        require(bytes % 16 == 0, "copy size must be a multiple of 16 bytes");
        const int64_t n = (int64_t)(bytes / 16);
        if (n == 0) return;
        const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream),
                           static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
        HIP_OK(hipGetLastError());
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
int gsmc_memcpy_htod_async(void* dst, const void* src, uint64_t bytes, void* stream)
{
    return guarded([&] {
        HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
    });
}
int gsmc_memcpy_dtoh_async(void* dst, const void* src, uint64_t bytes, void* stream)
{
    return guarded([&] {
        HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)));
    });
}
int gsmc_host_alloc(void** ptr, uint64_t bytes)
{
    return guarded([&] { HIP_OK(hipHostMalloc(ptr, std::max<uint64_t>(bytes, 1), hipHostMallocDefault)); });
}
int gsmc_host_free(void* ptr)
{
    return guarded([&] { HIP_OK(hipHostFree(ptr)); });
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
int gsmc_event_query(void* event, int* done)
{
    return guarded([&] {
        const hipError_t e = hipEventQuery(static_cast<hipEvent_t>(event));
        if (e == hipErrorNotReady) {
            *done = 0;
            return;
        }
        HIP_OK(e);
        *done = 1;
    });
}
int gsmc_event_elapsed_ms(void* start, void* stop, float* ms)
{
    return guarded([&] {
        HIP_OK(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
        HIP_OK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    });
}

}  // extern "C"
