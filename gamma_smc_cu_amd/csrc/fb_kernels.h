// Device code of the Gamma-SMC cached forward-backward decoder (gfx950).
//
// One lane per haplotype pair; a wave walks one work item -- 64 pairs of a
// job through the sites of one block window -- in order. Per block window
// (core [cs, ce) padded to [ps, pe)):
//   phase 1  forward filter over [ps, ce), storing the state at the start
//            of every K-site window of the core to HBM (pair-contiguous,
//            two windows per 16-byte record);
//   phase 2  backward filter over the right flank [ce, pe) (no output);
//   phase 3  windows of the core, last to first: recompute the K forward
//            states of the window from its checkpoint (into LDS at K = 8,
//            registers at K = 16), then walk the window backward,
//            combining forward and backward states into the posterior and
//            storing it pair-contiguous.
// Het calls come from XOR of the two bitpacked haplotype rows; the
// per-site transition is a bilinear lookup in the table layer selected by
// the site's bp gap, which is the same for every lane.
//
// Arithmetic follows src/kernels/gamma_smc_flow.cu operation for operation
// (bilinear :408-430, segments :471-514, fast 10^x :115-123, combine
// :708-752); the library is built with -ffp-contract=off, so the only
// fused multiply-adds are the explicit fmaf() calls below.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace gsmc {

constexpr int kGridNodes = 2550;
// Table slabs are stored by cell: cell (r, q), r <= 50, q <= 49, is 32
// bytes at (50 r + q) * 32, its entries (r, q) and (r, q + 1), each the
// 16-byte {m, c at node (r, q'); m, c at node (r + 1, q')}. A cell is the
// four corners one lookup reads, 32-byte aligned, so it never crosses a
// 128-byte cache line, and its offset is one multiply-add of the cell
// indices. The last row and column of cells (r = 50, q = 49) read node row
// 51 as a copy of row 49 and node column 50 as a copy of column 48; see
// lookup() for why that reproduces the reference at the grid's upper edges.
constexpr uint32_t kPairRows = 51;
constexpr uint32_t kRowCells = 50;
constexpr uint32_t kRowBytes = kRowCells * 32;
constexpr uint32_t kSlabBytes = kPairRows * kRowBytes;  // one slab: 81,600 bytes
constexpr uint32_t kLayerBytes = 2 * kSlabBytes;        // hom slab then het slab
constexpr uint32_t kDescLayerBits = 13;              // site descriptor: layer | chunks << 13
constexpr uint32_t kDescLayerMask = (1u << kDescLayerBits) - 1;

// Grid spacing as the reference's device code computes it
// (gamma_smc_flow.cu:10-17): float steps, then their float reciprocals.
constexpr float kMeanStep = (2.0f - -5.0f) / 50;
constexpr float kCvStep = (0.0f - -2.0f) / 49;
constexpr float kMeanStepInv = 1.0f / kMeanStep;
constexpr float kCvStepInv = 1.0f / kCvStep;

// Address-space views. Descriptors and launch metadata are read-only while a
// kernel runs, so they may go through the scalar (constant) cache; output
// and checkpoint pointers are global memory, which lets the compiler use
// a scalar base plus a per-lane offset instead of 64-bit flat addresses.
typedef __attribute__((address_space(4))) const uint32_t const_u32;
typedef __attribute__((address_space(1))) char global_char;

// Read-only metadata through the scalar cache, one dword at a time (the
// compiler merges the dwords into wide scalar loads).
template <class T> __device__ __forceinline__ T load_const(const T* p)
{
    static_assert(sizeof(T) % 4 == 0, "metadata records are dword multiples");
    T v;
    const const_u32* q = (const const_u32*)p;
    uint32_t* o = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) o[i] = q[i];
    return v;
}

template <class T> __device__ __forceinline__ T* global_at(void* base, uint32_t byte_off)
{
    return (T*)((global_char*)base + byte_off);
}

// Table bases as lookup() takes them: each the table's address less
// kFloatIntBias (see lookup()), never dereferenced without a biased offset.
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
    int64_t rows;      // sites of the job (rows of its summary partial planes)
    int64_t part_off;  // summary launches: element offset of the job's [groups][rows] partials
};

// One block window of a job. The walk starts from the prior at ps; the first
// site's forward/backward descriptors are sd_fwd / sd_bwd (the site's own gap
// descriptor at the start of a sequence, a 1-bp step at a block restart).
struct BlockDev {
    int32_t cs, ce, ps, pe;
    uint32_t sd_fwd, sd_bwd;
    int32_t job;
    int32_t nwin;    // K-site windows of the core
    int32_t q0;      // the walk starts from prior[q0]: the prior after q0 full N-bp hom stretches
    int32_t pad;
};

// A work item: one tile (64 pairs per lane slot) of one block window, and
// the float2 offset of its rows in the state buffer: one row of 64 NP
// states per core window (checkpoints) or per core site (stored forward
// states).
struct ItemDev {
    int32_t block;
    int32_t tile;
    int64_t ck_off;
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
    float2* ck;                // checkpoint records: per item, [ceil(nwin / 2)][64 NP] float4 (as float2 pairs)
    float2* yb;                // backward state at each item's core end: per item, [64 NP]
    int* queue2;               // persistent mode: item counter of the second kernel
    int64_t slot_elems;        // fused stored-state kernel: float2 state elements of each wave's slot
    DevTables t;
    float unscale;             // 2 Ne
    const float2* prior;       // prior[q] = q full N-bp hom stretches applied to the prior (0, 0)
    int32_t qbits;             // mean output encoding: 0 float, 8 / 4 log-quantized codes
    float q_inv_range;         // 1 / log(20 Ne): codes span [log 1, log 20 Ne]
    double* psum;              // summary launches: per 64-pair group and site, the sum of the means
    float* pmin;               // ... their minimum
    float* pmax;               // ... their maximum
    const int32_t* gene_piece = nullptr;  // gene launches: code of each site's piece (piece_of_code)
    double2* gene_part = nullptr;         // ... per piece and pair: (sum of x, sum of log2 x), x = max(mean, 1)
};

// Lane v of this lane's quad (DPP quad permutation; every lane active). The
// zero "old" value with bound_ctrl lets the compiler fold the permutation
// into the consuming add or multiply (one VOP2 DPP instruction instead of a
// move and the operation).
template <int CTRL> __device__ __forceinline__ uint32_t quad_perm(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}
template <int CTRL> __device__ __forceinline__ float quad_perm(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
constexpr int kQuad0101 = 0x44;  // lanes read lanes 0, 1, 0, 1 of their quad
constexpr int kQuad2323 = 0xEE;  // 2, 3, 2, 3
constexpr int kQuad2301 = 0x4E;  // 2, 3, 0, 1

// Byte offsets into the tables are formed by exact float multiply-adds
// from the cell indices; a float in [2^23, 2^24) holding an integer v has
// the bits kFloatIntBias + v, so the offset is read from those bits with
// the bias taken off the (scalar) table base instead of converting.
constexpr uint32_t kFloatIntBias = 0x4B000000u;  // bits of 2^23
constexpr uint32_t kHomSlab = kFloatIntBias;     // biased offset of a layer's hom slab
constexpr uint32_t kHetSlab = kFloatIntBias | kSlabBytes;

// Bilinear lookup of one table slab at (m, c); updates (m, c). layer: the
// layer's address less kFloatIntBias (DevTables); slab (kHomSlab or
// kHetSlab: the slab's byte offset, 0 or kSlabBytes, as biased float bits)
// selects the hom or het slab of the layer. Every lane of the wave must call
// it together.
//
// Same values as gamma_smc_flow.cu:408-430 in fewer instructions: the clamp
// is one med3 (identical for any non-NaN index) and the cell corner is
// floor(index) (truncation of a non-negative index), formed in float, exact
// for these small integers. The reference moves an index at the upper edge
// (50 for the mean, 49 for the CV) down one cell, with weight 1: corner row
// 49, value fma(1, A50, 0 * A49) where A_r interpolates node row r. Here
// that index stays in edge cell 50 with weight 0, whose node row 51 is a
// copy of row 49: fma(0, A49, 1 * A50). Both fused multiply-adds add the
// same two exact terms, A50 and a zero signed like A49, so the results are
// identical, zeros' signs and NaNs included; likewise for the CV column.
//
// Loads: the texture path handles a 16-byte gather four lanes at a time and
// pays one cache access per distinct line in each such quad. Each lane needs
// 32 contiguous bytes, its entries (r, q) and (r, q + 1), so the quad's
// first gather reads both for lanes 0 and 1 and the second both for lanes 2
// and 3 (two lines per quad and gather instead of up to four). A lane keeps
// the (r, q + 1) entry it loaded itself and receives its (r, q) entry from
// its partner lane (lane ^ 2) through a quad permutation that folds into the
// multiply by (1 - wc).
__device__ __forceinline__ void lookup(const char* __restrict__ layer, uint32_t slab, float& m, float& c)
{
    float fm = (m - -5.0f) * kMeanStepInv;
    float fc = (c - -2.0f) * kCvStepInv;
    fm = __builtin_amdgcn_fmed3f(fm, 0.0f, 50.0f);
    fc = __builtin_amdgcn_fmed3f(fc, 0.0f, 49.0f);
    const float lm = floorf(fm);
    const float lc = floorf(fc);
    const float wm = fm - lm;
    const float wc = fc - lc;
    // cell (lm, lc) at byte 32 (50 lm + lc) of the slab: biased float bits
    const uint32_t off = __float_as_uint(fmaf(lm, (float)kRowBytes, fmaf(lc, 32.0f, __uint_as_float(slab))));
    const bool upper = (__lane_id() & 2u) != 0;
    // first gather: lanes 0, 1 their entry (r, q + 1); lanes 2, 3 entry (r, q) of lanes 0, 1
    const uint32_t a1 = quad_perm<kQuad0101>(off) + (upper ? 0u : 16u);
    // second gather: lanes 2, 3 their entry (r, q + 1); lanes 0, 1 entry (r, q) of lanes 2, 3
    const uint32_t a2 = quad_perm<kQuad2323>(off) + (upper ? 16u : 0u);
    // scalar base (the table bases carry the bias) + 32-bit vector offsets
    const global_char* base = (const global_char*)layer;
    typedef float v4f __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) const v4f global_v4f;
    const v4f x = *reinterpret_cast<global_v4f*>(base + a1);
    const v4f y = *reinterpret_cast<global_v4f*>(base + a2);
    // own entry (r, q + 1) = {v01, v11}; the partner's copy of own entry (r, q) = {v00, v10}
    const v4f e1 = upper ? y : x;
    const v4f z = upper ? x : y;
    const float omwm = 1.0f - wm, omwc = 1.0f - wc;
    const float p00m = omwc * quad_perm<kQuad2301>(z.x), p10m = omwc * quad_perm<kQuad2301>(z.z);
    const float p00c = omwc * quad_perm<kQuad2301>(z.y), p10c = omwc * quad_perm<kQuad2301>(z.w);
    m = fmaf(wm, fmaf(wc, e1.z, p10m), omwm * fmaf(wc, e1.x, p00m));
    c = fmaf(wm, fmaf(wc, e1.w, p10c), omwm * fmaf(wc, e1.y, p00c));
}

__device__ __forceinline__ const char* layer_of(const char* base, uint32_t d)
{
    return base + (size_t)(d & kDescLayerMask) * kLayerBytes;
}

// Forward segment: full N-bp hom stretches, then the final table of
// (remaining bp, observed site).
__device__ __forceinline__ void fwd_segment(const DevTables& t, uint32_t d, uint32_t slab, float& m, float& c)
{
    for (uint32_t q = d >> kDescLayerBits; q != 0; --q) lookup(t.stretch, kHomSlab, m, c);
    lookup(layer_of(t.fwd, d), slab, m, c);
}

// Backward segment: observed site with the remaining bp first, then the
// full N-bp stretches.
__device__ __forceinline__ void bwd_segment(const DevTables& t, uint32_t d, uint32_t slab, float& m, float& c)
{
    lookup(layer_of(t.bwd, d), slab, m, c);
    for (uint32_t q = d >> kDescLayerBits; q != 0; --q) lookup(t.stretch, kHomSlab, m, c);
}

// Schraudolph-style 10^x as upstream gamma_smc reconstructs it
// (gamma_smc_flow.cu:115-123), bits = (int)(12102203.16 * (x * 2.30258509)
// + 1064872507.15) with the multiply-add fused as nvcc fuses it, given the
// product xk = x * 2.30258509299f.
__device__ __forceinline__ float pow10_fast_of_product(float xk)
{
    return __int_as_float((int)(fmaf(12102203.161561485f, xk, 1064872507.1541044f)));
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

// Bit k of a het mask, k a compile-time constant in the unrolled loops.
__device__ __forceinline__ uint32_t het_bit(uint64_t bits, int k)
{
    const uint32_t word = k < 32 ? (uint32_t)bits : (uint32_t)(bits >> 32);
    return (word >> (k & 31)) & 1u;
}

// The slab site k of a het mask selects, as lookup() takes it: kHomSlab or
// kHetSlab: a bit extract (0 or 1) and one 24-bit multiply-add, on 32
// bits. The extract is written out because the compiler otherwise rewrites
// it into a 64-bit compare and a select (four instructions).
__device__ __forceinline__ uint32_t het_slab(uint64_t bits, int k)
{
    const uint32_t word = k < 32 ? (uint32_t)bits : (uint32_t)(bits >> 32);
    uint32_t ext;
    asm("v_bfe_u32 %0, %1, %2, 1" : "=v"(ext) : "v"(word), "n"(k & 31));
    return __umul24(ext, kSlabBytes) + kFloatIntBias;
}

// K consecutive site descriptors from s (the array is padded by 64 entries
// past the last site), through scalar loads.
template <int K>
__device__ __forceinline__ void load_descs(const uint32_t* desc, int64_t s, uint32_t (&d)[K])
{
    const const_u32* p = (const const_u32*)(desc + s);
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = p[k];
}

// True when none of the K descriptors carries a full-stretch chunk (gap
// beyond the table depth): the window's lookups are then one table layer
// each and need no loop.
template <int K> __device__ __forceinline__ bool no_long_gaps(const uint32_t (&d)[K])
{
    uint32_t any = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) any |= d[k];
    return (any >> kDescLayerBits) == 0;
}

// Streaming store that bypasses cache retention (outputs are never re-read).
__device__ __forceinline__ void store_stream(float* row, uint32_t col, float v)
{
    __builtin_nontemporal_store(v, global_at<float>(row, col));
}

// Posterior mean of the forward (fm, fc) and backward (ym, yc) states
// (gamma_smc_flow.cu:708-723): alpha = 10^(-2 fc) + 10^(-2 yc) - 1, beta =
// 10^(-2 fc - fm) + 10^(-2 yc - ym) - 1. The reference forms -2 c first;
// that product is exact, so (-2 c) * k equals c * (-2 k) and (-2 c) - m
// equals fma(-2, c, -m), bit for bit, one instruction fewer each.
__device__ __forceinline__ float combine_mean(float fm, float fc, float ym, float yc, float unscale, float& as,
                                              float& bs)
{
    constexpr float k = 2.30258509299f, m2k = -2.0f * k;
    const float ba = pow10_fast_of_product(yc * m2k), bb = pow10_fast_of_product(fmaf(-2.0f, yc, -ym) * k);
    const float fa = pow10_fast_of_product(fc * m2k), fb = pow10_fast_of_product(fmaf(-2.0f, fc, -fm) * k);
    as = fa + ba - 1.0f;
    bs = fb + bb - 1.0f;
    return (as / fmaxf(bs, 1e-10f)) * unscale;
}

// The CI / posterior outputs of one site.
template <bool CI, bool POST>
__device__ __forceinline__ void store_extras(const JobDev& j, size_t row, uint32_t col, float mg, float as, float bs)
{
    if (CI) {
        // Wilson-Hilferty 95% bounds (gamma_smc_flow.cu:725-734).
        const float aci = fmaxf(as, 1.0f);
        const float inv9a = 1.0f / (9.0f * aci);
        const float sq = sqrtf(inv9a);
        const float base = 1.0f - inv9a;
        const float lo = fmaxf(fmaf(-1.96f, sq, base), 0.0f);
        const float hi = fmaf(1.96f, sq, base);
        store_stream(j.lower + row, col, fmaxf(mg * lo * lo * lo, 0.0f));
        store_stream(j.upper + row, col, mg * hi * hi * hi);
    }
    if (POST) {
        store_stream(j.alpha + row, col, as);
        store_stream(j.beta + row, col, bs);
    }
}

// Per-lane view of a work item: NP pairs per lane (lane l of the wave
// holds pairs tile * 64 NP + q * 64 + l, q < NP), their haplotype rows,
// output columns and checkpoint columns.
template <int NP> struct Lanes {
    const uint64_t* ri[NP];
    const uint64_t* rj[NP];
    uint32_t col[NP];   // byte offset of the pair in an output row
    bool live[NP];      // the pair exists (idle lanes shadow the job's last pair)
    uint32_t ck_col;    // byte offset of the lane in a row of 64 NP float2 (a record row is two such)
    float2* ck;         // state records of the item: record row t at ck + t * 128 NP
    int32_t grp0;       // 64-pair group of slot q = 0 (groups are tile * NP + q)
};

template <int K, int NP> struct Window {
    uint32_t d[K];     // site descriptors
    uint64_t bits[NP]; // het masks
};

// The het words X[w] = row_i[w] ^ row_j[w] around a walk's current window.
// A window's 64-site mask needs X[w] and X[w + 1] (w = first site / 64);
// windows of one walk step through the sites in order, so neighbouring
// windows share those words and the walk loads each word once (each load
// is a gather over 64 rows) instead of once per window.
template <int NP> struct HetCache {
    int64_t w = -4;           // word index of lo (uniform over the wave)
    uint64_t lo[NP], hi[NP];  // X[w], X[w + 1]; 0 past the end of the rows
};

template <int NP> __device__ __forceinline__ uint64_t het_word(const Lanes<NP>& L, int q, int64_t w, int64_t W)
{
    return (w < W) ? (L.ri[q][w] ^ L.rj[q][w]) : 0ull;
}

template <int K, int NP>
__device__ __forceinline__ void load_window(const FbArgs& a, const Lanes<NP>& L, int64_t ws, Window<K, NP>& win,
                                            HetCache<NP>& h)
{
    load_descs<K>(a.desc, ws, win.d);
    const int64_t w = ws >> 6;
    if (w == h.w + 1) {  // walking forward: slide up one word
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            h.lo[q] = h.hi[q];
            h.hi[q] = het_word(L, q, w + 1, a.W);
        }
    } else if (w == h.w - 1) {  // walking backward: slide down one word
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            h.hi[q] = h.lo[q];
            h.lo[q] = het_word(L, q, w, a.W);
        }
    } else if (w != h.w) {
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            h.lo[q] = het_word(L, q, w, a.W);
            h.hi[q] = het_word(L, q, w + 1, a.W);
        }
    }
    h.w = w;
    const int sh = (int)(ws & 63);
#pragma unroll
    for (int q = 0; q < NP; ++q) win.bits[q] = sh == 0 ? h.lo[q] : (h.lo[q] >> sh) | (h.hi[q] << (64 - sh));
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// State records: one 16-byte record per lane holds two states, records t
// of the item's 64 NP lanes in row t -- stored forward states of core
// sites (cs + 2t, cs + 2t + 1), or the checkpoints of windows 2t and
// 2t + 1. The address path spends the same cycles on a 16-byte access as
// on an 8-byte one, so pairing halves its share; and since the vector
// memory counter retires loads in order, each record load (an HBM round
// trip) holds up the table gathers issued after it, so one load per two
// checkpoint windows halves those stalls too. The records go out and come
// back once each, a kernel apart: streaming (non-temporal) accesses that
// leave the caches to the transition tables.
template <int NP> __device__ __forceinline__ f32x4* rec_at(const Lanes<NP>& L, int64_t t, int q)
{
    return global_at<f32x4>(L.ck + (size_t)t * 128 * NP, L.ck_col * 2u + (uint32_t)q * 1024u);
}
__device__ __forceinline__ void put_rec(f32x4* p, float m0, float c0, float m1, float c1)
{
    f32x4 v;
    v.x = m0;
    v.y = c0;
    v.z = m1;
    v.w = c1;
    __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ f32x4 get_rec(const f32x4* p) { return __builtin_nontemporal_load(p); }

// Float means of one full window, slot q, stored as row segments: each
// lane writes its K means down one column of an LDS tile [K][64], then
// each lane stores V consecutive pairs of one row (V = 2 or 4: 8- or
// 16-byte stores), K / V store instructions instead of K. The address
// path handles a store one 4-lane quad per cycle whatever its width, so
// wider stores cost it proportionally fewer cycles. Needs the whole slot
// live and rows aligned to V floats.
template <int K, int NP, int V>
__device__ __forceinline__ void store_means_rows(const JobDev& j, const Lanes<NP>& L, int q, int64_t ws,
                                                 const float (&col)[K], float* stage)
{
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int k = 0; k < K; ++k) stage[k * 64 + lane] = col[k];
    __syncthreads();
    constexpr uint32_t per_row = 64 / V;  // lanes per 64-pair row segment
    const uint32_t r = lane / per_row;    // first row of this lane (the wave covers V rows per store)
    const uint32_t c = (lane % per_row) * V;
    // byte offset of the lane's first pair: the slot's first pair, plus c
    const uint64_t lane_off = (uint64_t)(L.col[q] - lane * 4u + c * 4u) + (uint64_t)r * (uint64_t)j.P * 4u;
    global_char* base = (global_char*)j.mean + (uint64_t)(ws - j.row0) * (uint64_t)j.P * 4u + lane_off;
    const uint64_t stride = (uint64_t)V * (uint64_t)j.P * 4u;  // V rows
#pragma unroll
    for (int i = 0; i < K / V; ++i) {
        const float* src = stage + (r + i * V) * 64 + c;
        global_char* dst = base + (uint64_t)i * stride;
        if (V == 4)
            __builtin_nontemporal_store(*reinterpret_cast<const f32x4*>(src), (f32x4*)dst);
        else
            __builtin_nontemporal_store(*reinterpret_cast<const f32x2*>(src), (f32x2*)dst);
    }
    __syncthreads();  // the tile is rewritten by the next window
}

// Per-site reduction of one window's means over the 64 lanes of each pair
// group, for summary launches. The means go through LDS (one row of 65
// floats per site, so both the writes and the reads below are free of bank
// conflicts); lane (k = lane % 16, p = lane / 16) then sums lanes 16p ..
// 16p+15 of site k in order, in FP64 from 0.0, and a butterfly over p
// (partner p ^ 1, then p ^ 2) joins the four parts. Lanes without a pair
// hold 0 and are left out of min / max. The tree is the same for every K,
// so results do not depend on the checkpoint interval. One lane per site
// stores the group's partials.
template <int K, int NP>
__device__ __forceinline__ void summarize(const FbArgs& a, const JobDev& j, const Lanes<NP>& L, int64_t ws,
                                          int64_t cnt, const float (&out)[NP][K], float* red)
{
    const int lane = (int)threadIdx.x;
    // lanes 16 p .. 16 p + 15 read part p of rows 0..K-1 (k < K): within
    // each 32-lane half of a read, rows hit banks k and parts banks 16 p,
    // all distinct
    const int k = lane & 15, part = lane >> 4;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        // pairs of this group that exist: a prefix of the 64 lanes
        const int64_t first = (int64_t)(L.grp0 + q) * 64;
        const int live = (int)min(max(j.P - first, (int64_t)0), (int64_t)64);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < K; ++i) red[i * 65 + lane] = L.live[q] ? out[q][i] : 0.0f;
        __syncthreads();
        double sum = 0.0;
        float lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
        if (k < K) {
            const float* row = red + k * 65 + part * 16;
            if (live == 64) {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const float v = row[i];
                    sum = sum + (double)v;
                    lo = fminf(lo, v);
                    hi = fmaxf(hi, v);
                }
            } else {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const float v = row[i];
                    sum = sum + (double)v;
                    if (part * 16 + i < live) {
                        lo = fminf(lo, v);
                        hi = fmaxf(hi, v);
                    }
                }
            }
        }
#pragma unroll
        for (int mask = 16; mask <= 32; mask <<= 1) {  // parts p ^ 1, then p ^ 2
            sum = sum + __shfl_xor(sum, mask);
            lo = fminf(lo, __shfl_xor(lo, mask));
            hi = fmaxf(hi, __shfl_xor(hi, mask));
        }
        if (part == 0 && k < cnt) {
            const int64_t idx = j.part_off + (int64_t)(L.grp0 + q) * j.rows + (ws + k - j.row0);
            a.psum[idx] = sum;
            a.pmin[idx] = lo;
            a.pmax[idx] = hi;
        }
    }
}

// Log-quantized code of a posterior mean (gamma_smc.cu:23-39): the natural
// log of max(mean, 1 generation) mapped linearly from [0, log 20Ne] onto
// 0 .. levels and rounded half up (the scale-and-add is fused as nvcc
// fuses it), then clamped.
__device__ __forceinline__ uint32_t log_code(float mean, float inv_range, float levels)
{
    const float x = logf(fmaxf(mean, 1.0f)) * inv_range;
    const float q = fmaf(x, levels, 0.5f);
    return (uint32_t)fminf(fmaxf(q, 0.0f), levels);
}

// Compact output of one window: uint8 codes [rows][P], or uint4 codes two
// sites per byte ([ceil(rows / 2)][P], even row in the low nibble). Windows
// start at an even row (block cores have an even length), so no byte is
// shared between windows; an odd last row leaves the high nibble 0.
template <int K, int NP, bool FAST>
__device__ __forceinline__ void store_codes(const FbArgs& a, const JobDev& j, const Lanes<NP>& L, int64_t ws,
                                            int64_t cnt, const float (&out)[NP][K])
{
    global_char* base = (global_char*)j.mean;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        if (!L.live[q]) continue;
        const size_t col = L.col[q] >> 2;
        if (a.qbits == 8) {
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (FAST || k < cnt)
                    __builtin_nontemporal_store((uint8_t)log_code(out[q][k], a.q_inv_range, 255.0f),
                                                (uint8_t*)(base + (size_t)(ws + k - j.row0) * (size_t)j.P + col));
        } else {
#pragma unroll
            for (int k = 0; k < K; k += 2) {
                if (FAST || k < cnt) {
                    uint32_t v = log_code(out[q][k], a.q_inv_range, 15.0f);
                    if (FAST || k + 1 < cnt) v |= log_code(out[q][k + 1], a.q_inv_range, 15.0f) << 4;
                    __builtin_nontemporal_store((uint8_t)v,
                                                (uint8_t*)(base + (size_t)((ws + k - j.row0) >> 1) * (size_t)j.P + col));
                }
            }
        }
    }
}

// The forward states of the window being walked backward, slot s of pair
// q: in registers, or (LS) in LDS rows [NP][K][64] float2 of the wave,
// which keeps 2 K NP registers out of the backward kernel's allocation.
template <int K, int NP, bool LS> struct FwdStates {
    float m[NP][K], c[NP][K];
    __device__ __forceinline__ float2 get(int q, int s) const { return make_float2(m[q][s], c[q][s]); }
    __device__ __forceinline__ void set(int q, int s, float vm, float vc)
    {
        m[q][s] = vm;
        c[q][s] = vc;
    }
};
template <int K, int NP> struct FwdStates<K, NP, true> {
    float2* col;  // this lane's column of the rows
    __device__ __forceinline__ float2 get(int q, int s) const { return col[(q * K + s) * 64]; }
    __device__ __forceinline__ void set(int q, int s, float vm, float vc) { col[(q * K + s) * 64] = make_float2(vm, vc); }
};

// Per-gene epilogue: the lane's running sums over the sites of the current
// piece (a run of sites inside one block core and one elementary segment of
// the genes' boundaries; the same piece for the whole wave).
struct GeneAcc {
    double sl = 0.0, sg = 0.0;  // sum of x and of log2 x, x = max(mean, 1)
    int32_t piece = -1;
};

__device__ __forceinline__ void gene_flush(const FbArgs& a, const JobDev& j, uint32_t col, bool live, GeneAcc& G)
{
    if (G.piece >= 0 && live)
        a.gene_part[(size_t)G.piece * (size_t)j.P + (col >> 2)] = make_double2(G.sl, G.sg);
    G.sl = G.sg = 0.0;
}

// Site code: 2 k + 1 for a site of piece k, 2 k for a site in no gene with
// k pieces before it, so codes never decrease along the sites and a window
// whose first and last sites share a code lies inside one piece (or one gap).
__device__ __forceinline__ int32_t piece_of_code(int32_t code) { return (code & 1) ? (code >> 1) : -1; }

__device__ __forceinline__ void gene_site(const FbArgs& a, const JobDev& j, const Lanes<1>& L, int32_t pc, float mean,
                                          GeneAcc& G)
{
    if (pc != G.piece) {
        gene_flush(a, j, L.col[0], L.live[0], G);
        G.piece = pc;
    }
    if (pc >= 0) {
        const float x = fmaxf(mean, 1.0f);
        G.sl += (double)x;
        G.sg += (double)__log2f(x);
    }
}

// The window's sites in decreasing order: with the windows walked last to
// first this takes every core site in decreasing order, so the sums are the
// same for any checkpoint interval. A change of piece stores the finished
// piece's sums (a piece lies inside one block core, so exactly one item
// stores it per pair). Full windows inside one piece or gap take one test.
template <int K, bool FAST>
__device__ __forceinline__ void gene_accumulate(const FbArgs& a, const JobDev& j, const Lanes<1>& L, int64_t ws,
                                                int64_t cnt, const float (&out)[1][K], GeneAcc& G)
{
    if (FAST) {
        const int32_t c0 = __builtin_amdgcn_readfirstlane(a.gene_piece[ws]);
        const int32_t c1 = __builtin_amdgcn_readfirstlane(a.gene_piece[ws + K - 1]);
        if (c0 == c1) {
            const int32_t pc = piece_of_code(c0);
            if (pc != G.piece) {
                gene_flush(a, j, L.col[0], L.live[0], G);
                G.piece = pc;
            }
            if (pc >= 0) {
#pragma unroll
                for (int k = K - 1; k >= 0; --k) {
                    const float x = fmaxf(out[0][k], 1.0f);
                    G.sl += (double)x;
                    G.sg += (double)__log2f(x);
                }
            }
            return;
        }
    }
#pragma unroll
    for (int k = K - 1; k >= 0; --k)
        if (FAST || k < cnt)
            gene_site(a, j, L, piece_of_code(__builtin_amdgcn_readfirstlane(a.gene_piece[ws + k])), out[0][k], G);
}

// Phase-3 step over one checkpoint window, software-pipelined: the
// backward walk of window w (cnt sites at ws, forward states in F) runs in
// the same loop as the forward recompute of window w - 1 (K sites) when
// has_next. The two chains are independent, so each pair keeps two lookups
// in flight. The recomputed states reuse the slots the backward walk frees,
// which flips the slot order every window: orientation A holds site ws + k
// in slot k, orientation B in slot K-1-k. Window outputs are buffered and
// stored after the loop, so no lookup of the window waits on an output
// store (vector-memory counters are shared by loads and stores). FAST (a
// full window and a full next window, no long gaps in either) has no branch
// inside the loop, so the compiler can interleave all chains.
template <int K, int NP, bool CI, bool POST, bool SUM, bool B, bool FAST, bool ST, bool LS, bool GENE>
__device__ __forceinline__ void window_step(const FbArgs& a, const JobDev& j, const Lanes<NP>& L, int64_t cs,
                                            int64_t ws, int64_t cnt, bool has_next, const Window<K, NP>& cur,
                                            const Window<K, NP>& nxt, float (&m)[NP], float (&c)[NP],
                                            FwdStates<K, NP, LS>& F, float (&ym)[NP],
                                            float (&yc)[NP], float* red, GeneAcc& G)
{
    float out[NP][K], oa[NP][K], ob[NP][K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const int k = K - 1 - i;
        const int slot = B ? i : k;  // current window: site ws + k
        if (FAST || k < cnt) {
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const float2 f = F.get(q, slot);
                out[q][k] = combine_mean(f.x, f.y, ym[q], yc[q], a.unscale, oa[q][k], ob[q][k]);
                if (FAST)
                    lookup(layer_of(a.t.bwd, cur.d[k]), het_slab(cur.bits[q], k), ym[q], yc[q]);
                else
                    bwd_segment(a.t, cur.d[k], het_slab(cur.bits[q], k), ym[q], yc[q]);
            }
        } else if (SUM) {
#pragma unroll
            for (int q = 0; q < NP; ++q) out[q][k] = 0.0f;
        }
        if (FAST || has_next) {
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                if (ST) {
                    // next window: sites ws - K + i - 1 and ws - K + i, read back
                    // into the two slots just freed
                    if (i & 1) {
                        const f32x4 f = get_rec(rec_at(L, (ws - K + i - 1 - cs) >> 1, q));
                        const int prev = B ? i - 1 : K - i;
                        F.set(q, prev, f.x, f.y);
                        F.set(q, slot, f.z, f.w);
                    }
                    continue;
                }
                // the next window's checkpoint is already the state of its
                // first site (i = 0)
                if (i > 0) {
                    if (FAST)
                        lookup(layer_of(a.t.fwd, nxt.d[i]), het_slab(nxt.bits[q], i), m[q], c[q]);
                    else
                        fwd_segment(a.t, nxt.d[i], het_slab(nxt.bits[q], i), m[q], c[q]);
                }
                F.set(q, slot, m[q], c[q]);  // next window: site ws - K + i, in the slot just freed
            }
        }
        // Keep each site's work in its own scheduling region: the two chains
        // still interleave inside it, but nothing is hoisted from later
        // sites, which would cost registers and so waves.
        if (FAST) __builtin_amdgcn_sched_barrier(0);
    }
    if (SUM) {
        summarize<K, NP>(a, j, L, ws, cnt, out, red);
        return;
    }
    if constexpr (GENE) {
        static_assert(NP == 1, "gene launches decode one pair per lane");
        gene_accumulate<K, FAST>(a, j, L, ws, cnt, out, G);
        return;
    }
    if (!CI && !POST && a.qbits != 0) {
        store_codes<K, NP, FAST>(a, j, L, ws, cnt, out);
        return;
    }
    // row segments through LDS for full windows of full slots (float means
    // only; the stores below otherwise)
    const uintptr_t mb = reinterpret_cast<uintptr_t>(j.mean);
    const int vec = (!CI && !POST && FAST) ? (((j.P & 3) == 0 && (mb & 15) == 0) ? 4 : ((j.P & 1) == 0 && (mb & 7) == 0) ? 2 : 1)
                                          : 1;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        if (!CI && !POST && FAST && vec > 1 && (int64_t)(L.grp0 + q + 1) * 64 <= j.P) {
            if (vec == 4)
                store_means_rows<K, NP, 4>(j, L, q, ws, out[q], red);
            else
                store_means_rows<K, NP, 2>(j, L, q, ws, out[q], red);
            continue;
        }
        if (!L.live[q]) continue;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (FAST || k < cnt) {
                const size_t row = (size_t)(ws + k - j.row0) * (size_t)j.P;
                store_stream(j.mean + row, L.col[q], out[q][k]);
                if (CI || POST) store_extras<CI, POST>(j, row, L.col[q], out[q][k], oa[q][k], ob[q][k]);
            }
        }
    }
}

// The item's metadata and per-lane view, shared by both halves of a decode.
template <int NP> struct ItemView {
    BlockDev b;
    JobDev j;
    Lanes<NP> L;
    bool valid;
};

template <int NP>
__device__ __forceinline__ ItemView<NP> open_item(const FbArgs& a, int item_in, int lane, int slot)
{
    ItemView<NP> v;
    // The item is the same for the whole wave; saying so keeps every piece
    // of block and job metadata, and each site's descriptor, in scalar
    // registers and scalar loads.
    const int item = __builtin_amdgcn_readfirstlane(item_in);
    const ItemDev it = load_const(a.items + item);
    v.valid = it.block >= 0;  // padding of the XCD-ordered schedule
    if (!v.valid) return v;
    v.b = load_const(a.blocks + it.block);
    v.j = load_const(a.jobs + v.b.job);
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int64_t p = (int64_t)it.tile * 64 * NP + q * 64 + lane;
        v.L.live[q] = p < v.j.P;
        const int64_t pp = v.L.live[q] ? p : v.j.P - 1;
        v.L.col[q] = (uint32_t)p * 4u;
        // Pair ids from device memory are clamped to the cohort, so a bad id
        // can only produce a wrong row, never an out-of-bounds read; the host
        // entry points validate ids before they reach the device.
        const int64_t hi = min(max((int64_t)v.j.pair_i[pp], (int64_t)0), a.n_haps - 1);
        const int64_t hj = min(max((int64_t)v.j.pair_j[pp], (int64_t)0), a.n_haps - 1);
        v.L.ri[q] = a.packed + (size_t)hi * (size_t)a.W;
        v.L.rj[q] = a.packed + (size_t)hj * (size_t)a.W;
    }
    // Every lane, idle ones included, owns its state columns: the item's
    // rows of the launch's state buffer, or (slot >= 0) the rows of the
    // wave's own slot, reused by each item the wave decodes.
    v.L.ck = slot >= 0 ? a.ck + (size_t)slot * (size_t)a.slot_elems : a.ck + it.ck_off;
    v.L.ck_col = (uint32_t)lane * 8u;
    v.L.grp0 = it.tile * NP;
    return v;
}

template <int NP> __device__ __forceinline__ float2* yb_at(const FbArgs& a, int item, int lane, int q)
{
    return global_at<float2>(a.yb + (size_t)item * 64 * NP, (uint32_t)(q * 64 + lane) * 8u);
}

// First half of an item (64 NP pairs x one block window): the forward walk
// over [ps, ce) storing the state after the first site of every core window
// (ST: after every core site), and the backward walk over the right
// flank [ce, pe), whose final state is the backward state entering the core.
template <int K, int NP, bool ST>
__device__ __forceinline__ void forward_item(const FbArgs& a, int item, int lane, int slot = -1)
{
    const ItemView<NP> v = open_item<NP>(a, item, lane, slot);
    if (!v.valid) return;
    const BlockDev& b = v.b;
    const Lanes<NP>& L = v.L;
    const int64_t cs = b.cs, ce = b.ce, ps = b.ps, pe = b.pe;

    // Start state: the prior, advanced by q0 full stretches at a sequence
    // start (the same lookups the first site's segment would apply, done
    // once per context instead of once per item).
    const float2 x0 = load_const(a.prior + b.q0);
    float m[NP], c[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        m[q] = x0.x;
        c[q] = x0.y;
    }
    HetCache<NP> hf, hb;
    float ckm[NP], ckc[NP];  // the even window's checkpoint until its odd partner's
    // A core window's checkpoint is the state after its first site, so the
    // backward half recomputes K - 1 sites per window, not K. Checkpoints of
    // windows 2r and 2r + 1 share 16-byte record r; a last even window fills
    // both halves.
    auto checkpoint = [&](int64_t w, int q) {
        if (w & 1) {
            put_rec(rec_at(L, w >> 1, q), ckm[q], ckc[q], m[q], c[q]);
        } else if (w + 1 == b.nwin) {
            put_rec(rec_at(L, w >> 1, q), m[q], c[q], m[q], c[q]);
        } else {
            ckm[q] = m[q];
            ckc[q] = c[q];
        }
    };
    for (int64_t s = ps; s < ce;) {
        const int64_t e = (s < cs) ? min(cs, s + K) : min(ce, s + K);
        const bool core = s >= cs;  // windows are all flank or all core
        const int64_t w = (s - cs) / K;
        Window<K, NP> win;
        load_window(a, L, s, win, hf);
        if (s == ps) win.d[0] = b.sd_fwd;
        // ST: core windows start an even number of sites after cs, so site
        // pairs (k - 1, k), k odd, are records; a window's odd last site
        // fills both halves of its record.
        float pm[NP], pc[NP];
        if (e - s == K && no_long_gaps<K>(win.d)) {
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    lookup(layer_of(a.t.fwd, win.d[k]), het_slab(win.bits[q], k), m[q], c[q]);
                    if (!ST && core && k == 0) checkpoint(w, q);
                    if (ST && core) {
                        if (k & 1) {
                            put_rec(rec_at(L, (s - cs + k) >> 1, q), pm[q], pc[q], m[q], c[q]);
                        } else {
                            pm[q] = m[q];
                            pc[q] = c[q];
                        }
                    }
                }
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (s + k < e)
#pragma unroll
                    for (int q = 0; q < NP; ++q) {
                        fwd_segment(a.t, win.d[k], het_slab(win.bits[q], k), m[q], c[q]);
                        if (!ST && core && k == 0) checkpoint(w, q);
                        if (ST && core) {
                            if (k & 1) {
                                put_rec(rec_at(L, (s - cs + k) >> 1, q), pm[q], pc[q], m[q], c[q]);
                            } else if (s + k + 1 == e) {
                                put_rec(rec_at(L, (s - cs + k) >> 1, q), m[q], c[q], m[q], c[q]);
                            } else {
                                pm[q] = m[q];
                                pc[q] = c[q];
                            }
                        }
                    }
        }
        s = e;
    }

    float ym[NP], yc[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) ym[q] = yc[q] = 0.0f;
    for (int64_t e = pe; e > ce;) {
        const int64_t b0 = max(ce, e - K);
        Window<K, NP> win;
        load_window(a, L, b0, win, hb);
        if (e - b0 == K && no_long_gaps<K>(win.d)) {
#pragma unroll
            for (int k = K - 1; k >= 0; --k)
#pragma unroll
                for (int q = 0; q < NP; ++q) lookup(layer_of(a.t.bwd, win.d[k]), het_slab(win.bits[q], k), ym[q], yc[q]);
        } else {
#pragma unroll
            for (int k = K - 1; k >= 0; --k)
                if (b0 + k < e)
#pragma unroll
                    for (int q = 0; q < NP; ++q) bwd_segment(a.t, win.d[k], het_slab(win.bits[q], k), ym[q], yc[q]);
        }
        e = b0;
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) *yb_at<NP>(a, __builtin_amdgcn_readfirstlane(item), lane, q) = make_float2(ym[q], yc[q]);
}

// Second half of an item: the core windows, last to first. The last window
// (possibly partial) is recomputed alone into orientation A; every later
// recompute rides along the previous window's backward walk. ST: the
// forward states are read back instead of recomputed, so the next
// window's descriptors and het bits are not needed.
template <int K, int NP, bool CI, bool POST, bool SUM, bool ST, bool LS = false, bool GENE = false>
__device__ __forceinline__ void backward_item(const FbArgs& a, int item, int lane, float* red, int slot = -1,
                                              float2* states = nullptr)
{
    const ItemView<NP> v = open_item<NP>(a, item, lane, slot);
    if (!v.valid) return;
    const BlockDev& b = v.b;
    const JobDev& j = v.j;
    const Lanes<NP>& L = v.L;
    const int64_t cs = b.cs, ce = b.ce, ps = b.ps;

    float m[NP], c[NP], ym[NP], yc[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const float2 y = *yb_at<NP>(a, __builtin_amdgcn_readfirstlane(item), lane, q);
        ym[q] = y.x;
        yc[q] = y.y;
        m[q] = c[q] = 0.0f;
    }
    FwdStates<K, NP, LS> F;
    if constexpr (LS) F.col = states + lane;
    float ckm[NP], ckc[NP];  // a loaded record's other checkpoint (window w - 1 when w is odd)
    int64_t w = (int64_t)b.nwin - 1;
    int64_t cnt = ce - (cs + w * K);
    HetCache<NP> hc, hn;  // current windows; next windows (recomputed forward states)
    {
        const int64_t ws = cs + w * K;
        if (ST) {
#pragma unroll
            for (int k = 0; k < K; k += 2) {
                if (k < cnt) {
#pragma unroll
                    for (int q = 0; q < NP; ++q) {
                        const f32x4 f = get_rec(rec_at(L, (ws + k - cs) >> 1, q));
                        F.set(q, k, f.x, f.y);
                        F.set(q, k + 1, f.z, f.w);
                    }
                }
            }
        } else {
            Window<K, NP> win;
            load_window(a, L, ws, win, hn);  // descriptor 0 unused: site ws is the checkpoint
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                // record w >> 1 holds the checkpoints of windows w & ~1 and w | 1
                const f32x4 r = get_rec(rec_at(L, w >> 1, q));
                m[q] = (w & 1) ? r.z : r.x;
                c[q] = (w & 1) ? r.w : r.y;
                ckm[q] = r.x;  // window w - 1's checkpoint when w is odd
                ckc[q] = r.y;
                F.set(q, 0, m[q], c[q]);  // the checkpoint is the state of site ws
            }
#pragma unroll
            for (int k = 1; k < K; ++k) {
                if (k < cnt) {
#pragma unroll
                    for (int q = 0; q < NP; ++q) {
                        fwd_segment(a.t, win.d[k], het_slab(win.bits[q], k), m[q], c[q]);
                        F.set(q, k, m[q], c[q]);
                    }
                }
            }
        }
    }
    bool orient_b = false;
    GeneAcc G;
    for (;;) {
        const bool has_next = w > 0;
        const int64_t ws = cs + w * K;
        Window<K, NP> cur, nxt;
        load_window(a, L, ws, cur, hc);
        if (ws == ps) cur.d[0] = b.sd_bwd;
        if (has_next && !ST) {
            load_window(a, L, ws - K, nxt, hn);
            nxt.d[0] = 0;  // unused: site ws - K is the checkpoint (and no long gap)
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                if ((w - 1) & 1) {
                    // one load for windows w - 1 (now) and w - 2 (next window)
                    const f32x4 r = get_rec(rec_at(L, (w - 1) >> 1, q));
                    m[q] = r.z;
                    c[q] = r.w;
                    ckm[q] = r.x;
                    ckc[q] = r.y;
                } else {
                    m[q] = ckm[q];
                    c[q] = ckc[q];
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) nxt.d[k] = 0;
#pragma unroll
            for (int q = 0; q < NP; ++q) nxt.bits[q] = 0;
        }
        const bool fast = has_next && cnt == K && no_long_gaps<K>(cur.d) && (ST || no_long_gaps<K>(nxt.d));
        if (fast) {
            if (orient_b)
                window_step<K, NP, CI, POST, SUM, true, true, ST, LS, GENE>(a, j, L, cs, ws, cnt, true, cur, nxt, m, c, F,
                                                                            ym, yc, red, G);
            else
                window_step<K, NP, CI, POST, SUM, false, true, ST, LS, GENE>(a, j, L, cs, ws, cnt, true, cur, nxt, m, c,
                                                                             F, ym, yc, red, G);
        } else {
            if (orient_b)
                window_step<K, NP, CI, POST, SUM, true, false, ST, LS, GENE>(a, j, L, cs, ws, cnt, has_next, cur, nxt, m,
                                                                             c, F, ym, yc, red, G);
            else
                window_step<K, NP, CI, POST, SUM, false, false, ST, LS, GENE>(a, j, L, cs, ws, cnt, has_next, cur, nxt,
                                                                              m, c, F, ym, yc, red, G);
        }
        if (!has_next) break;
        orient_b = !orient_b;
        --w;
        cnt = K;
    }
    if constexpr (GENE) gene_flush(a, j, L.col[0], L.live[0], G);
}


// Forward states of the backward kernels: in LDS for K = 8 (recomputed
// states), which takes the K = 8 mean kernel from 123 to under 80 registers, so
// six waves per SIMD instead of four hide the latency of each lookup's
// gathers; in registers for K = 16 (its 8 KB of states per wave would cap
// the waves per CU) and for stored states.
template <int K, bool ST> constexpr bool kStatesInLds = !ST && K == 8;
template <int K, int NP, bool ST> constexpr int kStateLds = kStatesInLds<K, ST> ? K * NP * 64 : 1;

// Waves per SIMD the backward kernel is built for (a floor for the register
// allocator): five with one pair per lane and forward states in LDS (the
// mean kernel needs under 80 registers and gets six; the persistent and CI /
// posterior kernels are held to 96 registers, five waves), three with them
// in registers (a few registers over would leave two), four for stored
// states with two pairs per lane (two lookup chains per wave; 134 registers
// unbounded, three waves), no floor otherwise.
template <int K, int NP, bool ST>
constexpr int kBwdWaves = (NP == 1 && kStatesInLds<K, ST>) ? 5 : (NP == 1 && !ST) ? 3 : (NP == 2 && ST) ? 4 : 1;

// LDS floats of the backward kernel: the summary's staging rows (one row
// of 65 per site of the window), or the [K][64] tile of float-mean row
// stores (reused slot by slot).
template <int K, bool CI, bool POST, bool SUM>
constexpr int kBwdLds = SUM ? K * 65 : (CI || POST) ? 1 : K * 64;

// A decode is two launches over the same work items: fb_fwd_* (forward
// walk + right-flank backward walk; few registers, so many waves per SIMD
// to hide the latency of its single lookup chain) then fb_bwd_* (the
// pipelined core windows). Static schedule: workgroup g decodes item g;
// persistent schedule: waves take items from a device counter until it
// passes n_items, so every wave exits.
template <int K, int NP, bool ST>
__global__ __launch_bounds__(64) void fb_fwd_items(FbArgs a)
{
    const int item = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
    if (item < a.n_items) forward_item<K, NP, ST>(a, item, threadIdx.x);
}

template <int K, int NP, bool CI, bool POST, bool SUM, bool ST>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kBwdWaves<K, NP, ST>))) void fb_bwd_items(FbArgs a)
{
    const int item = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
    __shared__ float red[kBwdLds<K, CI, POST, SUM>];
    __shared__ float2 states[kStateLds<K, NP, ST>];
    if (item < a.n_items)
        backward_item<K, NP, CI, POST, SUM, ST, kStatesInLds<K, ST>>(a, item, threadIdx.x, red, -1, states);
}

// Per-gene launches: the backward walk of fb_bwd_items with the gene
// epilogue in place of row stores (one pair per lane, static schedule).
template <int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kBwdWaves<K, 1, false>))) void fb_bwd_genes(
    FbArgs a)
{
    const int item = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
    __shared__ float red[1];
    __shared__ float2 states[kStateLds<K, 1, false>];
    if (item < a.n_items)
        backward_item<K, 1, false, false, false, false, kStatesInLds<K, false>, true>(a, item, threadIdx.x, red, -1,
                                                                                      states);
}

// Lane 0 takes the next item of the counter; every caller reaches this with
// all 64 lanes active (uniform control flow), so the first active lane is
// lane 0 and readfirstlane broadcasts its value without a lane permute (the
// permute kept extra registers live and cost the backward kernels a wave).
__device__ __forceinline__ int next_item(int* counter, int lane)
{
    int item = 0;
    if (lane == 0) item = atomicAdd(counter, 1);
    return __builtin_amdgcn_readfirstlane(item);
}

// Dynamic schedule: one workgroup per item, as the static grid, but each
// workgroup decodes the next item of the device counter (items in their
// longest-first order) instead of item blockIdx.x, so items are handed to
// whichever XCD frees a slot first; no item loop, so no values live across
// items.
template <int K, int NP, bool ST>
__global__ __launch_bounds__(64) void fb_fwd_grab(FbArgs a)
{
    const int item = next_item(a.queue, threadIdx.x);
    if (item < a.n_items) forward_item<K, NP, ST>(a, item, threadIdx.x);
}

template <int K, int NP, bool CI, bool POST, bool SUM, bool ST>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kBwdWaves<K, NP, ST>))) void fb_bwd_grab(FbArgs a)
{
    __shared__ float red[kBwdLds<K, CI, POST, SUM>];
    __shared__ float2 states[kStateLds<K, NP, ST>];
    const int item = next_item(a.queue2, threadIdx.x);
    if (item < a.n_items)
        backward_item<K, NP, CI, POST, SUM, ST, kStatesInLds<K, ST>>(a, item, threadIdx.x, red, -1, states);
}

template <int K, int NP, bool ST>
__global__ __launch_bounds__(64) void fb_fwd_queue(FbArgs a)
{
    for (int item = next_item(a.queue, threadIdx.x); item < a.n_items; item = next_item(a.queue, threadIdx.x))
        forward_item<K, NP, ST>(a, item, threadIdx.x);
}

template <int K, int NP, bool CI, bool POST, bool SUM, bool ST>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kBwdWaves<K, NP, ST>))) void fb_bwd_queue(FbArgs a)
{
    __shared__ float red[kBwdLds<K, CI, POST, SUM>];
    __shared__ float2 states[kStateLds<K, NP, ST>];
    for (int item = next_item(a.queue2, threadIdx.x); item < a.n_items; item = next_item(a.queue2, threadIdx.x))
        backward_item<K, NP, CI, POST, SUM, ST, kStatesInLds<K, ST>>(a, item, threadIdx.x, red, -1, states);
}

// Stored forward states, one launch: persistent waves take items from the
// device counter and decode each whole -- the forward half into the wave's
// slot of the state buffer, then the backward half reading it back -- so
// the state buffer holds one item per resident wave, not one per item, and
// one pool of waves drains all items (no second kernel waiting on the
// first one's last wave). The fence makes the slot's states and the
// backward entry state written by the forward half visible to the
// backward half's loads.
template <int KF, int KB, int NP, bool SUM, int WAVES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVES))) void fb_fused_queue(FbArgs a)
{
    __shared__ float red[kBwdLds<KB, false, false, SUM>];
    const int slot = (int)blockIdx.x;
    for (int item = next_item(a.queue, threadIdx.x); item < a.n_items; item = next_item(a.queue, threadIdx.x)) {
        forward_item<KF, NP, true>(a, item, threadIdx.x, slot);
        __threadfence();
        backward_item<KB, NP, false, false, SUM, true>(a, item, threadIdx.x, red, slot);
    }
}

// Gather-only probe of the texture path: every lane runs two independent
// chains of the decode's table lookup on the context's forward tables, one
// layer per step shared by the wave (drawn from a hash over the N layers)
// and the het slab in about 15% of lanes, with nothing else in the loop.
// Its lookups per second bound what the decode's lookups can reach on this
// device. Writes one float per lane so the chains are kept.
__device__ __forceinline__ uint32_t probe_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(64) void lookup_probe_kernel(DevTables t, int32_t n_layers, int32_t steps, float* out)
{
    const uint32_t gid = blockIdx.x * 64u + threadIdx.x;
    float m[2], c[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const uint32_t h = probe_hash(gid * 7919u + (uint32_t)k * 104729u + 17u);
        m[k] = -1.5f + 2.5f * (float)(h & 0xFFFFu) * (1.0f / 65535.0f);
        c[k] = -1.2f + 0.9f * (float)(h >> 16) * (1.0f / 65535.0f);
    }
    for (int32_t r = 0; r < steps; ++r) {
        const uint32_t lay = __builtin_amdgcn_readfirstlane(probe_hash((uint32_t)r * 31u + blockIdx.x) % (uint32_t)n_layers);
        const char* layer = t.fwd + (size_t)lay * kLayerBytes;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint32_t hs = probe_hash(gid * 13u + (uint32_t)r * 7u + (uint32_t)k);
            lookup(layer, (hs & 1023u) < 159u ? kHetSlab : kHomSlab, m[k], c[k]);
        }
    }
    out[gid] = m[0] + c[0] + m[1] + c[1];
}

// Launch metadata of one blockwise job, built on the device (for the entry
// point that allocates nothing): the job, block windows b < nb of `core`
// sites with `flank`-site flanks over [0, S), and items b * tiles + t with
// a uniform state stride. A window starting at site 0 starts the sequence
// (descriptor sd0 from prior[q0]); any other starts with descriptor
// sd_restart (one bp). Same values as the host plan for this job.
__global__ __launch_bounds__(256) void build_meta_kernel(JobDev job, JobDev* jobs, BlockDev* blocks, ItemDev* items,
                                                         int64_t S, int64_t core, int64_t flank, int32_t nb,
                                                         int32_t tiles, uint32_t sd0, int32_t q0, uint32_t sd_restart,
                                                         int32_t K, int64_t ck_stride)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) jobs[0] = job;
    if (i < nb) {
        const int64_t cs = i * core, ce = min(S, cs + core);
        BlockDev d;
        d.cs = (int32_t)cs;
        d.ce = (int32_t)ce;
        d.ps = (int32_t)max((int64_t)0, cs - flank);
        d.pe = (int32_t)min(S, ce + flank);
        d.sd_fwd = d.ps == 0 ? sd0 : sd_restart;
        d.sd_bwd = 0;
        d.job = 0;
        d.nwin = (int32_t)((ce - cs + K - 1) / K);
        d.q0 = d.ps == 0 ? q0 : 0;
        d.pad = 0;
        blocks[i] = d;
    }
    if (i < (int64_t)nb * tiles) {
        ItemDev it;
        it.block = (int32_t)(i / tiles);
        it.tile = (int32_t)(i % tiles);
        it.ck_off = i * ck_stride;
        items[i] = it;
    }
}

// Forward states of the full-sequence walk (the reference's run_fwd_states
// debug path, bindings.cpp:2851-2893 over gamma_smc_flow.cu:567-609): one
// lane per pair, the decode's forward segments from the prior, and the
// state (m, c) after every site written site-major to out_m / out_c
// [S][P]. The first site starts from prior[q0] and applies descriptor sd0
// (its gap from position -1, full stretches folded into q0).
__global__ __launch_bounds__(64) void fwd_states_kernel(FbArgs a, const int32_t* __restrict__ pi,
                                                        const int32_t* __restrict__ pj, int64_t P, int64_t S,
                                                        int32_t q0, uint32_t sd0, float* out_m, float* out_c)
{
    const int lane = (int)threadIdx.x;
    const int64_t p = (int64_t)blockIdx.x * 64 + lane;
    const bool live = p < P;
    const int64_t pp = live ? p : P - 1;
    const int64_t hi = min(max((int64_t)pi[pp], (int64_t)0), a.n_haps - 1);
    const int64_t hj = min(max((int64_t)pj[pp], (int64_t)0), a.n_haps - 1);
    Lanes<1> L;
    L.ri[0] = a.packed + (size_t)hi * (size_t)a.W;
    L.rj[0] = a.packed + (size_t)hj * (size_t)a.W;
    const float2 x0 = load_const(a.prior + q0);
    float m = x0.x, c = x0.y;
    HetCache<1> h;
    for (int64_t s = 0; s < S; s += 16) {
        Window<16, 1> win;
        load_window(a, L, s, win, h);
        if (s == 0) win.d[0] = sd0;
        const int64_t cnt = min((int64_t)16, S - s);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k < cnt) {
                fwd_segment(a.t, win.d[k], het_slab(win.bits[0], k), m, c);
                if (live) {
                    out_m[(size_t)(s + k) * (size_t)P + (size_t)p] = m;
                    out_c[(size_t)(s + k) * (size_t)P + (size_t)p] = c;
                }
            }
        }
    }
}

// Summary epilogue: per site r of a job, the groups' partials in group
// order -- the FP64 sum g = 0, 1, ... (starting from the running total of
// earlier pair chunks unless `first`), min and max. The last chunk writes
// mean = (float)(sum / n_pairs), min and max; earlier chunks keep the
// running values in tot / tmin / tmax.
__global__ __launch_bounds__(256) void summary_reduce_kernel(const double* __restrict__ psum,
                                                             const float* __restrict__ pmin,
                                                             const float* __restrict__ pmax, int64_t groups,
                                                             int64_t rows, double* tot, float* tmin, float* tmax,
                                                             int first, int last, double n_pairs, float* site_mean,
                                                             float* site_min, float* site_max)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double t = first ? 0.0 : tot[r];
    float lo = first ? __builtin_huge_valf() : tmin[r];
    float hi = first ? -__builtin_huge_valf() : tmax[r];
    for (int64_t g = 0; g < groups; ++g) {
        t = t + psum[g * rows + r];
        lo = fminf(lo, pmin[g * rows + r]);
        hi = fmaxf(hi, pmax[g * rows + r]);
    }
    if (last) {
        site_mean[r] = (float)(t / n_pairs);
        if (site_min) site_min[r] = lo;
        if (site_max) site_max[r] = hi;
    } else {
        tot[r] = t;
        tmin[r] = lo;
        tmax[r] = hi;
    }
}

// prior[q] for q < n: q applications of the N-bp hom stretch to the prior
// (0, 0), one dependent lookup each, with the decode kernels' own lookup.
// (one wave: every lane walks the same chain, lookups need the whole wave;
// lane 0 stores)
__global__ void prior_table_kernel(DevTables t, float2* out, int64_t n)
{
    if (blockIdx.x != 0) return;
    const bool writer = threadIdx.x == 0;
    float m = 0.0f, c = 0.0f;
    if (writer) out[0] = make_float2(m, c);
    for (int64_t q = 1; q < n; ++q) {
        lookup(t.stretch, kHomSlab, m, c);
        if (writer) out[q] = make_float2(m, c);
    }
}

// ---------------------------------------------------------------------------
// Forward-only filter (run_fwd): the reference's half-precision cached path,
// gamma_smc_flow.cu:1139-1257. Tables: the composed transitions cache[k]
// (N layers of 2550 (m, c) nodes) rounded to FP16 pairs. Per site: advance
// by the site's gap (descriptor kNoAdvance at site 0 and for gaps that round
// to 0; full N-bp layers first), apply the het emission explicitly, output
// 10^m * 2Ne (and the Wilson-Hilferty bounds). One lane per pair, so a
// launch has only P / 64 waves and its time is the per-site latency of the
// dependent chain times the sites: one 16-byte gather per lookup and the
// fast 10^x keep that chain short.
constexpr uint32_t kNoAdvance = 0xFFFFFFFFu;

// FP16 slab layout: entry (r, q) (r <= 50, q <= 49) holds the four corners
// (r, q), (r + 1, q), (r, q + 1), (r + 1, q + 1) as FP16 (m, c) pairs, 16
// bytes at 16 (50 r + q), so a lookup is one aligned gather; node row 51
// repeats row 49 and node column 50 repeats column 48 (the edge cells of
// the FP32 tables, so the index needs no cap). 40,800 bytes per layer.
constexpr uint32_t kQuadSlabEntries = 51 * 50;

struct FwdArgs {
    const uint64_t* packed;
    int64_t n_haps;
    int64_t W;
    int64_t S;
    const uint32_t* desc;   // [S + 64] forward-only gap descriptors
    const uint4* h2q;       // [N][51 * 50] quad entries
    int32_t N;
    const int32_t* pair_i;
    const int32_t* pair_j;
    int64_t P;
    float two_ne;
    float* mean;            // [S][P]
    float* lower;
    float* upper;
};

__device__ __forceinline__ float2 half2_bits(uint32_t v)
{
    return __half22float2(*reinterpret_cast<const __half2*>(&v));
}

// Bilinear lookup on one FP16 slab, the arithmetic of the FP32 lookup
// (gamma_smc_flow.cu:1139-1172). The walk's time is the latency of its
// chain of dependent instructions per site, so the entry's byte offset is
// one multiply-add from the mean index: 16 (50 lm + lc) + 2^23, an exact
// integer in float whose bits are the offset plus 0x4B000000 (taken off the
// layer base instead of converting), with no cap on the index (edge cells).
__device__ __forceinline__ void lookup_h2(const uint4* __restrict__ layer, float& m, float& c)
{
    float fm = (m - -5.0f) * kMeanStepInv;
    float fc = (c - -2.0f) * kCvStepInv;
    fm = __builtin_amdgcn_fmed3f(fm, 0.0f, 50.0f);
    fc = __builtin_amdgcn_fmed3f(fc, 0.0f, 49.0f);
    const float lm = floorf(fm), lc = floorf(fc);
    const float wm = fm - lm, wc = fc - lc;
    // scalar layer base (less the bias) + 32-bit vector offset
    const uint32_t off = __float_as_uint(fmaf(lm, 800.0f, fmaf(lc, 16.0f, 8388608.0f)));
    const uint4 e = *global_at<const uint4>((void*)((uintptr_t)layer - kFloatIntBias), off);
    const float2 v00 = half2_bits(e.x), v10 = half2_bits(e.y), v01 = half2_bits(e.z), v11 = half2_bits(e.w);
    const float omwm = 1.0f - wm, omwc = 1.0f - wc;
    m = fmaf(wm, fmaf(wc, v11.x, omwc * v10.x), omwm * fmaf(wc, v01.x, omwc * v00.x));
    c = fmaf(wm, fmaf(wc, v11.y, omwc * v10.y), omwm * fmaf(wc, v01.y, omwc * v00.y));
}

template <bool CI>
__global__ __launch_bounds__(64) void fwd_only_kernel(FwdArgs a)
{
    const int64_t p = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const bool live = p < a.P;
    const int64_t pp = live ? p : a.P - 1;
    const int64_t hi = min(max((int64_t)a.pair_i[pp], (int64_t)0), a.n_haps - 1);
    const int64_t hj = min(max((int64_t)a.pair_j[pp], (int64_t)0), a.n_haps - 1);
    Lanes<1> L;
    L.ri[0] = a.packed + (size_t)hi * (size_t)a.W;
    L.rj[0] = a.packed + (size_t)hj * (size_t)a.W;
    HetCache<1> h;
    const uint4* last = a.h2q + (size_t)(a.N - 1) * kQuadSlabEntries;
    float m = 0.0f, c = 0.0f;
    for (int64_t s0 = 0; s0 < a.S; s0 += 16) {
        uint32_t d[16];
        load_descs<16>(a.desc, s0, d);
        // het words as the decode kernels cache them (one word pair per 64 sites)
        const int64_t w = s0 >> 6;
        if (w != h.w) {
            h.lo[0] = het_word(L, 0, w, a.W);
            h.hi[0] = het_word(L, 0, w + 1, a.W);
            h.w = w;
        }
        const int sh = (int)(s0 & 63);
        const uint64_t bits = sh == 0 ? h.lo[0] : (h.lo[0] >> sh) | (h.hi[0] << (64 - sh));
        const int cnt = (int)min((int64_t)16, a.S - s0);
        float out[16], olo[16], ohi[16];
        // PLAIN: the site has one transition of at most N bp (no stretch
        // chunks, not a site without a transition)
        auto site = [&](int k, auto plain) {
            const uint32_t dk = d[k];
            if (decltype(plain)::value) {
                lookup_h2(a.h2q + (size_t)dk * kQuadSlabEntries, m, c);
            } else if (dk != kNoAdvance) {
                for (uint32_t q = dk >> kDescLayerBits; q != 0; --q) lookup_h2(last, m, c);
                lookup_h2(a.h2q + (size_t)(dk & kDescLayerMask) * kQuadSlabEntries, m, c);
            }
            {
                // het emission, fast 10^x and log10 where the reference uses
                // its fast intrinsics (gamma_smc_flow.cu:1234-1235): log10 as
                // the hardware log2 times log10(2) (alpha >= 1, so no
                // denormal range to scale for). Some lane of the wave is het
                // at almost every site, so every lane computes it and the
                // het lanes take it (no divergent branch per site).
                const bool het = het_bit(bits, k);
                // 10^(-2c) as 2^(c * -2 log2(10)); -log10(alpha) / 2 with the
                // halving folded into the constant (both exact rescalings)
                const float alpha = __builtin_amdgcn_exp2f(c * -6.64385618977472470f) + 1.0f;
                const float l2 = __builtin_amdgcn_logf(alpha);
                const float b_log = fmaf(-2.0f, c, -m);
                // log10(alpha) - b_log in one fused step
                m = het ? fmaf(l2, 0.30102999566398120f, -b_log) : m;
                c = het ? l2 * -0.15051499783199060f : c;
            }
            // 10^m by the hardware 2^x: table states lie in the grid box
            // (m in [-5, 2]) and an emission raises m by at most log10(2),
            // far from the range ends that need scaling
            const float mg = __builtin_amdgcn_exp2f(m * 3.321928094887362f) * a.two_ne;
            out[k] = mg;
            if (CI) {
                const float alpha = __exp10f(-2.0f * c);
                const float inv9a = 1.0f / (9.0f * alpha);
                const float sq = sqrtf(inv9a);
                const float base = 1.0f - inv9a;
                const float lo = fmaxf(fmaf(-1.96f, sq, base), 0.0f);
                const float hf = fmaf(1.96f, sq, base);
                olo[k] = fmaxf(mg * lo * lo * lo, 0.0f);
                ohi[k] = mg * hf * hf * hf;
            }
        };
        auto store = [&](int k) {
            const size_t idx = (size_t)(s0 + k) * (size_t)a.P + (size_t)p;
            __builtin_nontemporal_store(out[k], a.mean + idx);
            if (CI) {
                __builtin_nontemporal_store(olo[k], a.lower + idx);
                __builtin_nontemporal_store(ohi[k], a.upper + idx);
            }
        };
        // Full chunks run without a bounds test per site, and, when every
        // site is plain (the usual case; the no-transition marker has its
        // high bits set too), without a descriptor test per site.
        uint32_t any_high = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) any_high |= d[k];
        if (cnt == 16 && (any_high >> kDescLayerBits) == 0) {
#pragma unroll
            for (int k = 0; k < 16; ++k) site(k, std::true_type{});
            if (live) {
#pragma unroll
                for (int k = 0; k < 16; ++k) store(k);
            }
        } else if (cnt == 16) {
#pragma unroll
            for (int k = 0; k < 16; ++k) site(k, std::false_type{});
            if (live) {
#pragma unroll
                for (int k = 0; k < 16; ++k) store(k);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (k < cnt) site(k, std::false_type{});
            if (live) {
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    if (k < cnt) store(k);
            }
        }
    }
}

// Node-major slabs ([n_slabs][2550] (m, c) nodes, mean-major) to the
// 32-byte cells the lookups read (kSlabBytes per slab): entry e = 2 (50 r +
// q) + h of a slab holds nodes (r, q + h) and (r + 1, q + h), node row 51
// reading row 49 and node column 50 reading column 48 (the edge cells).
__global__ void cells_from_nodes_kernel(const float2* __restrict__ nodes, float4* __restrict__ cells, int64_t n_slabs)
{
    constexpr int kEntries = (int)(kSlabBytes / 16);  // 5,100
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_slabs * kEntries) return;
    const int64_t s = i / kEntries;
    const int e = (int)(i - s * kEntries);
    const int r = e / 100, col = (e % 100) / 2 + (e & 1);
    const int q = col == 50 ? 48 : col, r1 = r == 50 ? 49 : r + 1;
    const float2* slab = nodes + s * kGridNodes;
    const float2 lo = slab[r * 50 + q], hi = slab[r1 * 50 + q];
    cells[i] = make_float4(lo.x, lo.y, hi.x, hi.y);
}

// FP32 (m, c) nodes [N][51][50] to FP16 quad entries [N][50][50] (round to
// nearest even, as the reference rounds its table with __floats2half2_rn):
// entry (r, q) = nodes (r, q), (r + 1, q), (r, q + 1), (r + 1, q + 1); the
// q + 1 of the unused last column repeats q.
__global__ void to_half2_quads_kernel(const float2* __restrict__ src, uint4* __restrict__ dst, int64_t n_layers)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_layers * (int64_t)kQuadSlabEntries) return;
    const int64_t layer = i / kQuadSlabEntries;
    const int e = (int)(i - layer * kQuadSlabEntries);
    const int r = e / 50, q = e % 50;
    const int r1 = r < 50 ? r + 1 : 49, q1 = q < 49 ? q + 1 : 48;
    const float2* nodes = src + layer * kGridNodes;
    auto bits = [](float2 v) {
        const __half2 h = __floats2half2_rn(v.x, v.y);
        return *reinterpret_cast<const uint32_t*>(&h);
    };
    dst[i] = make_uint4(bits(nodes[r * 50 + q]), bits(nodes[r1 * 50 + q]), bits(nodes[r * 50 + q1]),
                        bits(nodes[r1 * 50 + q1]));
}

__global__ void reset_counters_kernel(int* counters)
{
    counters[0] = 0;
    counters[1] = 0;
}

// Plain streaming copy (16 B per lane per step), used to calibrate HBM
// traffic counters against a kernel of known size.
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

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
