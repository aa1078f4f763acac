// Test infrastructure: a C entry point around the reference's own host table
// builder, so the oracle's FP32 restatement can be compared with it bit for
// bit. This file holds only glue written for this repository; the code it
// calls is compiled from /root/reference/src/flow_field.cpp where it lies
// (see the Makefile beside this file) and never copied.
#include "gamma_smc_cu/flow_field.h"
#include <cstring>

extern "C" int ref_build_tables(const char* path, int n_steps, float scaled_rho,
                                float scaled_mu, float* out, float* bounds)
{
    FlowFieldData ff;
    if (!load_flow_field(path, ff)) return -1;
    bounds[0] = ff.mean_log10_min;
    bounds[1] = ff.mean_log10_max;
    bounds[2] = ff.cv_log10_min;
    bounds[3] = ff.cv_log10_max;
    FlowFieldCache c = build_flow_field_cache(ff, n_steps, scaled_rho, scaled_mu);
    const float* order[10] = {c.mean, c.cv, c.fwd_hom_site_mean, c.fwd_hom_site_cv,
                              c.fwd_het_site_mean, c.fwd_het_site_cv, c.bwd_hom_site_mean,
                              c.bwd_hom_site_cv, c.bwd_het_site_mean, c.bwd_het_site_cv};
    size_t L = (size_t)n_steps * FF_GRID;
    for (int k = 0; k < 10; k++) std::memcpy(out + k * L, order[k], L * sizeof(float));
    free_flow_field_cache(c);
    return 0;
}

extern "C" int ref_load_flow_field(const char* path, float* u, float* v, float* bounds)
{
    FlowFieldData ff;
    if (!load_flow_field(path, ff)) return -1;
    std::memcpy(u, ff.u, sizeof ff.u);
    std::memcpy(v, ff.v, sizeof ff.v);
    bounds[0] = ff.mean_log10_min;
    bounds[1] = ff.mean_log10_max;
    bounds[2] = ff.cv_log10_min;
    bounds[3] = ff.cv_log10_max;
    return 0;
}
