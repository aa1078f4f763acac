// Device code of the per-gene statistics of posterior means (gfx950).
//
// The production caller of the decoder (analysis/genome_wide/
// infer_chromosome.py:157-227) decodes pair chunks to host and reduces the
// (sites, pairs) means there, per gene: for each pair the mean over the
// gene's sites of max(mean, 1) and of its natural log, then over the pairs
// the count, the sums of both and of the squared log, their minima and a
// histogram of the per-pair log. Here the means never leave the decode:
// the backward kernel's gene epilogue (fb_bwd_genes, fb_kernels.h) sums, per
// pair and piece (a run of sites inside one block core and one elementary
// segment of the genes' boundaries), x = max(mean, 1) and log2 x in FP64,
// and gene_reduce_kernel / gene_combine_kernel below turn the piece sums
// into the per-gene accumulators: per gene, the pieces it covers summed per
// pair in piece order, the pairs reduced in splits (several workgroups per
// gene, so that few genes still fill the chip) by a fixed tree, and the
// splits added in order to the running totals of earlier pair chunks. The
// reduction reads 16 B per pair and covered piece.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsmc {

constexpr int kMaxGeneBins = 1024;  // histogram bins held in LDS per gene

struct GeneReduceArgs {
    const double2* part;       // [n_pieces][P] piece sums of the gene epilogue
    int64_t P;
    const int32_t* gene_pb;    // gene g covers pieces [gene_pb[g], gene_pe[g])
    const int32_t* gene_pe;
    const int64_t* gene_n;     // sites of gene g
    const double* edges;       // n_bins + 1 increasing histogram edges (natural log)
    int n_bins;
    int64_t* count;            // running totals, one per gene
    double* lin_sum;
    double* log_sum;
    double* log_sq_sum;
    double* min_lin;
    double* min_log;
    int64_t* hist;             // [n_genes][n_bins]
    int32_t n_split;           // workgroups per gene; split s takes pairs [s P / n_split, (s + 1) P / n_split)
    double* split_val;         // [n_genes][n_split][5]: the split's three sums and two minima
    uint32_t* split_hist;      // [n_genes][n_split][n_bins]
};

// Workgroup (g, s) reduces split s of the pairs for gene g. Thread t takes
// the split's pairs t, t + 256, ...: the pair's piece sums in piece order
// give its mean x and mean log x (log2 sum times ln 2, over the gene's n
// sites); the thread keeps FP64 sums and minima, and the bin of mean log x
// (the number of edges <= it, less one, clamped to the bins: numpy.digitize
// and clip) is counted in LDS. A fixed tree over the 256 threads joins the
// sums and minima into the split's partials; gene_combine_kernel adds the
// splits in order to the running totals. Genes of fewer than 2 sites are
// left unchanged.
__global__ __launch_bounds__(256) void gene_reduce_kernel(GeneReduceArgs a)
{
    constexpr double kLn2 = 0.69314718055994530942;
    const int g = blockIdx.x, sp = blockIdx.y;
    const int t = threadIdx.x;
    const int64_t n = a.gene_n[g];
    if (n < 2) return;
    __shared__ double s_edge[kMaxGeneBins + 1];
    __shared__ uint32_t s_hist[kMaxGeneBins];
    __shared__ double s_red[5][256];
    for (int i = t; i <= a.n_bins; i += 256) s_edge[i] = a.edges[i];
    for (int i = t; i < a.n_bins; i += 256) s_hist[i] = 0u;
    __syncthreads();
    const int32_t pb = a.gene_pb[g], pe = a.gene_pe[g];
    const int64_t p0 = a.P * sp / a.n_split, p1 = a.P * (sp + 1) / a.n_split;
    const double inv_n = 1.0 / (double)n;
    double sl = 0.0, sg = 0.0, sq = 0.0;
    double ml = __builtin_inf(), mg = __builtin_inf();
    for (int64_t p = p0 + t; p < p1; p += 256) {
        double xl = 0.0, xg = 0.0;
        for (int32_t k = pb; k < pe; ++k) {
            const double2 d = a.part[(size_t)k * (size_t)a.P + (size_t)p];
            xl += d.x;
            xg += d.y;
        }
        const double lin = xl * inv_n;
        const double lg = (xg * kLn2) * inv_n;
        sl += lin;
        sg += lg;
        sq += lg * lg;
        ml = fmin(ml, lin);
        mg = fmin(mg, lg);
        int lo = 0, hi = a.n_bins + 1;  // count of edges <= lg, by bisection
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s_edge[mid] <= lg)
                lo = mid + 1;
            else
                hi = mid;
        }
        const int bin = min(max(lo - 1, 0), a.n_bins - 1);
        atomicAdd(&s_hist[bin], 1u);
    }
    s_red[0][t] = sl;
    s_red[1][t] = sg;
    s_red[2][t] = sq;
    s_red[3][t] = ml;
    s_red[4][t] = mg;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s_red[0][t] += s_red[0][t + w];
            s_red[1][t] += s_red[1][t + w];
            s_red[2][t] += s_red[2][t + w];
            s_red[3][t] = fmin(s_red[3][t], s_red[3][t + w]);
            s_red[4][t] = fmin(s_red[4][t], s_red[4][t + w]);
        }
        __syncthreads();
    }
    const size_t at = (size_t)g * (size_t)a.n_split + (size_t)sp;
    if (t < 5) a.split_val[at * 5 + t] = s_red[t][0];
    for (int i = t; i < a.n_bins; i += 256) a.split_hist[at * (size_t)a.n_bins + i] = s_hist[i];
}

// Per gene: the splits' partials in split order added to the running
// totals (count += P for the launch's pairs).
__global__ __launch_bounds__(256) void gene_combine_kernel(GeneReduceArgs a)
{
    const int g = blockIdx.x;
    const int t = threadIdx.x;
    if (a.gene_n[g] < 2) return;
    const size_t at = (size_t)g * (size_t)a.n_split;
    if (t == 0) {
        double sl = a.lin_sum[g], sg = a.log_sum[g], sq = a.log_sq_sum[g], ml = a.min_lin[g], mg = a.min_log[g];
        for (int s = 0; s < a.n_split; ++s) {
            const double* v = a.split_val + (at + s) * 5;
            sl += v[0];
            sg += v[1];
            sq += v[2];
            ml = fmin(ml, v[3]);
            mg = fmin(mg, v[4]);
        }
        a.count[g] += a.P;
        a.lin_sum[g] = sl;
        a.log_sum[g] = sg;
        a.log_sq_sum[g] = sq;
        a.min_lin[g] = ml;
        a.min_log[g] = mg;
    }
    for (int i = t; i < a.n_bins; i += 256) {
        int64_t h = 0;
        for (int s = 0; s < a.n_split; ++s) h += a.split_hist[(at + s) * (size_t)a.n_bins + i];
        a.hist[(size_t)g * (size_t)a.n_bins + i] += h;
    }
}

}  // namespace gsmc
