// Constants shared by the dedispersion kernel (dedisp.hip) and its host-side
// LDS planner (dedisp_plan.cpp).
#pragma once

namespace dsx {

constexpr int DD_CHUNK = 64;      // centres per candidate chunk
constexpr int DD_SUB = 32;        // centres per register pass (two per chunk)
constexpr int DD_THREADS = 256;   // DMs per work-group (grid.y covers the rest)
constexpr int DD_MAX_ROWS = 4096;
constexpr int DD_STAGE_ITERS = 12;   // fast path: a subband window spans at most 12 x 64 steps
constexpr int DD_STAGE_SUBS = 4;     // fast path: subbands a wave stages per batch of loads

// LDS words per thread holding its term bases (matches the kernel's MAXT choice).
inline int tb_words(int n_terms) { return n_terms <= 8 ? 4 : n_terms <= 16 ? 8 : 0; }

}  // namespace dsx
