// Test stand-in for green::utils::context() (single process).
#pragma once
namespace green::utils {
  struct context_t { int node_rank = 0, node_size = 1, global_rank = 0, global_size = 1; };
  inline const context_t& context() { static context_t c; return c; }
}  // namespace green::utils
