// Device code of the Gamma-SMC cached forward-backward decoder (gfx950).
//
// One lane per haplotype pair; a wave walks one work item -- 64 pairs of a
// job through the sites of one block window -- in order. Per block window
// (core [cs, ce) padded to [ps, pe)):
//   phase 1  forward filter over [ps, ce), storing the state at the start
//            of every K-site window of the core to HBM (pair-contiguous);
//   phase 2  backward filter over the right flank [ce, pe) (no output);
//   phase 3  windows of the core, last to first: recompute the K forward
//            states of the window from its checkpoint into registers, then
//            walk the window backward, combining forward and backward
//            states into the posterior and storing it pair-contiguous.
// Het calls come from XOR of the two bitpacked haplotype rows; the
// per-site transition is a bilinear lookup in the table layer selected by
// the site's bp gap, which is the same for every lane.
//
// Arithmetic follows src/kernels/gamma_smc_flow.cu operation for operation
// (bilinear :408-430, segments :471-514, fast 10^x :115-123, combine
// :708-752); the library is built with -ffp-contract=off, so the only
// fused multiply-adds are the explicit fmaf() calls below.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsmc {

constexpr int kGridNodes = 2550;
constexpr uint32_t kSlabBytes = kGridNodes * 8;      // one [2550] (m, c) float2 slab
constexpr uint32_t kLayerBytes = 2 * kSlabBytes;     // hom slab then het slab
constexpr uint32_t kDescLayerBits = 13;              // site descriptor: layer | chunks << 13
constexpr uint32_t kDescLayerMask = (1u << kDescLayerBits) - 1;

// Grid spacing as the reference's device code computes it
// (gamma_smc_flow.cu:10-17): float steps, then their float reciprocals.
constexpr float kMeanStep = (2.0f - -5.0f) / 50;
constexpr float kCvStep = (0.0f - -2.0f) / 49;
constexpr float kMeanStepInv = 1.0f / kMeanStep;
constexpr float kCvStepInv = 1.0f / kCvStep;

struct DevTables {
    const char* fwd;      // [N] layers of kLayerBytes
    const char* bwd;      // [N] layers of kLayerBytes
    const char* stretch;  // one slab: N-bp hom stretch
};

// One decode job: a set of pairs over a site range of the cohort, with
// its own outputs [rows][P] (row r = site row0 + r).
struct JobDev {
    float* mean;
    float* lower;
    float* upper;
    float* alpha;
    float* beta;
    const int32_t* pair_i;
    const int32_t* pair_j;
    int64_t P;
    int64_t row0;
};

// One block window of a job. The walk starts from the prior at ps; the first
// site's forward/backward descriptors are sd_fwd / sd_bwd (the site's own gap
// descriptor at the start of a sequence, a 1-bp step at a block restart).
struct BlockDev {
    int32_t cs, ce, ps, pe;
    uint32_t sd_fwd, sd_bwd;
    int32_t job;
    int32_t nwin;    // checkpoint windows of the core
    int64_t ck_off;  // float2 offset of this block's checkpoints (tile-major)
};

// A work item: one 64-pair tile of one block window.
struct ItemDev {
    int32_t block;
    int32_t tile;
};

struct FbArgs {
    const uint64_t* packed;    // [n_haps][W] bitpacked rows
    int64_t n_haps;
    int64_t W;
    const uint32_t* desc;      // [S] site descriptors (gap from the previous site)
    const JobDev* jobs;
    const BlockDev* blocks;
    const ItemDev* items;
    int64_t n_items;
    int* queue;                // persistent mode: next item counter (starts at 0)
    float2* ck;                // checkpoints: per item, [nwin][64]
    DevTables t;
    float unscale;             // 2 Ne
};

// Bilinear lookup of one table slab at (m, c); updates (m, c).
// het_off selects the hom (0) or het (kSlabBytes) slab of a layer.
__device__ __forceinline__ void lookup(const char* __restrict__ layer, uint32_t het_off, float& m, float& c)
{
    float fm = (m - -5.0f) * kMeanStepInv;
    float fc = (c - -2.0f) * kCvStepInv;
    fm = fmaxf(0.0f, fminf(fm, 50.0f));
    fc = fmaxf(0.0f, fminf(fc, 49.0f));
    const int m0 = min((int)fm, 49);
    const int c0 = min((int)fc, 48);
    const float wm = fm - (float)m0;
    const float wc = fc - (float)c0;
    const uint32_t off = het_off + (uint32_t)(m0 * 50 + c0) * 8u;
    const float2 v00 = *reinterpret_cast<const float2*>(layer + off);
    const float2 v01 = *reinterpret_cast<const float2*>(layer + off + 8);
    const float2 v10 = *reinterpret_cast<const float2*>(layer + off + 400);
    const float2 v11 = *reinterpret_cast<const float2*>(layer + off + 408);
    const float omwm = 1.0f - wm, omwc = 1.0f - wc;
    m = fmaf(wm, fmaf(wc, v11.x, omwc * v10.x), omwm * fmaf(wc, v01.x, omwc * v00.x));
    c = fmaf(wm, fmaf(wc, v11.y, omwc * v10.y), omwm * fmaf(wc, v01.y, omwc * v00.y));
}

// Forward segment: full N-bp hom stretches, then the final table of
// (remaining bp, observed site).
__device__ __forceinline__ void fwd_segment(const DevTables& t, uint32_t d, uint32_t het_off, float& m, float& c)
{
    for (uint32_t q = d >> kDescLayerBits; q != 0; --q) lookup(t.stretch, 0, m, c);
    lookup(t.fwd + (size_t)(d & kDescLayerMask) * kLayerBytes, het_off, m, c);
}

// Backward segment: observed site with the remaining bp first, then the
// full N-bp stretches.
__device__ __forceinline__ void bwd_segment(const DevTables& t, uint32_t d, uint32_t het_off, float& m, float& c)
{
    lookup(t.bwd + (size_t)(d & kDescLayerMask) * kLayerBytes, het_off, m, c);
    for (uint32_t q = d >> kDescLayerBits; q != 0; --q) lookup(t.stretch, 0, m, c);
}

// Schraudolph-style 10^x as upstream gamma_smc reconstructs it
// (gamma_smc_flow.cu:115-123); the multiply-add is fused as nvcc fuses it.
__device__ __forceinline__ float pow10_fast(float x)
{
    const int bits = (int)(fmaf(12102203.161561485f, x * 2.30258509299f, 1064872507.1541044f));
    return __int_as_float(bits);
}

// Het bits of the 64 sites starting at s0 (bit k = site s0 + k).
__device__ __forceinline__ uint64_t het_mask(const uint64_t* __restrict__ ri, const uint64_t* __restrict__ rj,
                                             int64_t s0, int64_t W)
{
    const int64_t w = s0 >> 6;
    const int sh = (int)(s0 & 63);
    const uint64_t x0 = ri[w] ^ rj[w];
    if (sh == 0) return x0;
    const uint64_t x1 = (w + 1 < W) ? (ri[w + 1] ^ rj[w + 1]) : 0ull;
    return (x0 >> sh) | (x1 << (64 - sh));
}

__device__ __forceinline__ uint32_t het_off_of(uint64_t bits, int k)
{
    return ((uint32_t)(bits >> k) & 1u) * kSlabBytes;
}

template <bool CI, bool POST>
__device__ __forceinline__ void combine_store(const JobDev& j, float unscale, int64_t s, int64_t p, bool live,
                                              float fm, float fc, float ym, float yc)
{
    const float lb = -2.0f * yc;
    const float ba = pow10_fast(lb), bb = pow10_fast(lb - ym);
    const float lf = -2.0f * fc;
    const float fa = pow10_fast(lf), fb = pow10_fast(lf - fm);
    const float as = fa + ba - 1.0f;
    const float bs = fb + bb - 1.0f;
    const float mg = (as / fmaxf(bs, 1e-10f)) * unscale;
    if (!live) return;
    const size_t idx = (size_t)(s - j.row0) * (size_t)j.P + (size_t)p;
    j.mean[idx] = mg;
    if (CI) {
        // Wilson-Hilferty 95% bounds (gamma_smc_flow.cu:725-734).
        const float aci = fmaxf(as, 1.0f);
        const float inv9a = 1.0f / (9.0f * aci);
        const float sq = sqrtf(inv9a);
        const float base = 1.0f - inv9a;
        const float lo = fmaxf(fmaf(-1.96f, sq, base), 0.0f);
        const float hi = fmaf(1.96f, sq, base);
        j.lower[idx] = fmaxf(mg * lo * lo * lo, 0.0f);
        j.upper[idx] = mg * hi * hi * hi;
    }
    if (POST) {
        j.alpha[idx] = as;
        j.beta[idx] = bs;
    }
}

// Decode one work item (64 pairs x one block window) with the calling wave.
template <int K, bool CI, bool POST>
__device__ __forceinline__ void decode_item(const FbArgs& a, int64_t item, int lane)
{
    const ItemDev it = a.items[item];
    if (it.block < 0) return;  // padding of the XCD-ordered schedule
    const BlockDev b = a.blocks[it.block];
    const JobDev j = a.jobs[b.job];
    const int64_t p = (int64_t)it.tile * 64 + lane;
    const bool live = p < j.P;
    const int64_t pp = live ? p : j.P - 1;  // idle lanes shadow the job's last pair
    const int64_t cs = b.cs, ce = b.ce, ps = b.ps, pe = b.pe;
    // Pair ids from device memory are clamped to the cohort, so a bad id can
    // only produce a wrong row, never an out-of-bounds read; the host entry
    // points validate ids before they reach the device.
    const int64_t hi = min(max((int64_t)j.pair_i[pp], (int64_t)0), a.n_haps - 1);
    const int64_t hj = min(max((int64_t)j.pair_j[pp], (int64_t)0), a.n_haps - 1);
    const uint64_t* ri = a.packed + (size_t)hi * (size_t)a.W;
    const uint64_t* rj = a.packed + (size_t)hj * (size_t)a.W;
    // Every lane, idle ones included, owns a checkpoint column.
    float2* ck = a.ck + (size_t)b.ck_off + (size_t)it.tile * (size_t)b.nwin * 64 + lane;

    // Phase 1: forward over [ps, ce), checkpoints at the core windows.
    float m = 0.0f, c = 0.0f;
    for (int64_t s = ps; s < ce;) {
        const int64_t e = (s < cs) ? min(cs, s + K) : min(ce, s + K);
        if (s >= cs) ck[(size_t)((s - cs) / K) * 64] = make_float2(m, c);
        const uint64_t bits = het_mask(ri, rj, s, a.W);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (s + k < e) {
                const uint32_t d = (k == 0 && s == ps) ? b.sd_fwd : a.desc[s + k];
                fwd_segment(a.t, d, het_off_of(bits, k), m, c);
            }
        }
        s = e;
    }

    // Phase 2: backward over the right flank [ce, pe).
    float ym = 0.0f, yc = 0.0f;
    for (int64_t e = pe; e > ce;) {
        const int64_t b0 = max(ce, e - K);
        const uint64_t bits = het_mask(ri, rj, b0, a.W);
#pragma unroll
        for (int k = K - 1; k >= 0; --k) {
            if (b0 + k < e) bwd_segment(a.t, a.desc[b0 + k], het_off_of(bits, k), ym, yc);
        }
        e = b0;
    }

    // Phase 3: core windows, last to first.
    for (int64_t w = (int64_t)b.nwin - 1; w >= 0; --w) {
        const int64_t ws = cs + w * K;
        const int64_t we = min(ce, ws + K);
        const float2 st = ck[(size_t)w * 64];
        m = st.x;
        c = st.y;
        const uint64_t bits = het_mask(ri, rj, ws, a.W);
        float fm[K], fc[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (ws + k < we) {
                const uint32_t d = (k == 0 && ws == ps) ? b.sd_fwd : a.desc[ws + k];
                fwd_segment(a.t, d, het_off_of(bits, k), m, c);
                fm[k] = m;
                fc[k] = c;
            }
        }
#pragma unroll
        for (int k = K - 1; k >= 0; --k) {
            if (ws + k < we) {
                combine_store<CI, POST>(j, a.unscale, ws + k, p, live, fm[k], fc[k], ym, yc);
                const uint32_t d = (k == 0 && ws == ps) ? b.sd_bwd : a.desc[ws + k];
                bwd_segment(a.t, d, het_off_of(bits, k), ym, yc);
            }
        }
    }
}

// Static schedule: wave w of workgroup g decodes item g * (WG / 64) + w.
template <int K, bool CI, bool POST, int WG>
__global__ __launch_bounds__(WG) void fb_items_kernel(FbArgs a)
{
    const int lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * (WG / 64) + (threadIdx.x >> 6);
    if (item < a.n_items) decode_item<K, CI, POST>(a, item, lane);
}

// Persistent schedule: each wave takes items from a device counter until
// the queue is drained; every wave exits once the counter passes n_items.
template <int K, bool CI, bool POST, int WG>
__global__ __launch_bounds__(WG) void fb_queue_kernel(FbArgs a)
{
    const int lane = threadIdx.x & 63;
    for (;;) {
        int item = 0;
        if (lane == 0) item = atomicAdd(a.queue, 1);
        item = __shfl(item, 0);
        if (item >= a.n_items) break;
        decode_item<K, CI, POST>(a, item, lane);
    }
}

__global__ void reset_counter_kernel(int* counter) { *counter = 0; }

// Bitpack: one wave per (haplotype, 64-site word); the ballot of the 64
// lanes' bits is the word (bit b of word w = G[h][64 w + b] & 1, tail 0).
__global__ __launch_bounds__(256) void bitpack_kernel(const uint8_t* __restrict__ G, int64_t n, int64_t S,
                                                      int64_t W, uint64_t* __restrict__ packed)
{
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t total = n * W;
    const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t t = wave; t < total; t += stride) {
        const int64_t h = t / W, w = t - h * W;
        const int64_t s = w * 64 + lane;
        const bool bit = (s < S) && (G[h * S + s] & 1);
        const uint64_t word = __ballot(bit);
        if (lane == 0) packed[t] = word;
    }
}

__global__ __launch_bounds__(256) void unpack_kernel(const uint64_t* __restrict__ packed, int64_t n, int64_t S,
                                                     int64_t W, uint8_t* __restrict__ G)
{
    const int64_t total = n * S;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t h = i / S, s = i - h * S;
        G[i] = (uint8_t)((packed[h * W + (s >> 6)] >> (s & 63)) & 1ull);
    }
}

}  // namespace gsmc
