// Host-side flow-field parsing and transition-table construction.
#pragma once

#include <string>
#include <vector>

namespace gsmc {

constexpr int kMeanN = 51;            // log10 mean grid, [-5, 2]
constexpr int kCvN = 50;              // log10 CV grid, [-2, 0]
constexpr int kGrid = kMeanN * kCvN;  // 2550 nodes, mean-major

enum { kTablesF32 = 0, kTablesF64 = 1 };

struct FlowField {
    float u[kGrid];
    float v[kGrid];
    float bounds[4];  // log10 of the header's linear bounds: mean lo, hi, cv lo, hi
};

struct HostTables {
    int n_steps = 0;
    std::vector<float> fwd;           // [N][2 (hom, het)][kGrid] x (m, c)
    std::vector<float> bwd;           // [N][2 (hom, het)][kGrid] x (m, c)
    std::vector<float> stretch_last;  // [kGrid] x (m, c): N-bp hom stretch
    std::vector<float> cache;         // [N][kGrid] x (m, c): composed k+1-bp transition (forward-only path)
};

// Parse the text format of src/flow_field.cpp:8-54 (two header lines, then
// 2550 u values and 2550 v values, mean-major).
bool load_flow_field(const char* path, FlowField& ff, std::string& err);

// scaled_rho / scaled_mu are the FP32 values 4.0f*Ne*rho, 4.0f*Ne*mu used by
// the FP32 build; the *64 values are their FP64 counterparts.
void build_tables(const FlowField& ff, int n_steps, float scaled_rho, float scaled_mu,
                  double scaled_rho64, double scaled_mu64, int precision, HostTables& out);

}  // namespace gsmc
