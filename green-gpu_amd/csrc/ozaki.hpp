// Complex GEMM with FP64 operands emulated on the gfx950 int8 matrix cores
// (Compute::I8X5 / I8X4): operands sliced into five or four signed 7-bit
// digits per element with a power-of-two scale per row of A / column of B,
// exact int32 slice products up to slice diagonal 5 / 4, FP64 recombination.
#pragma once

#include "zgemm.hpp"

namespace ggpu {

// Enqueue d (complex128 operands; complex128 or complex64 output) on the
// int8 matrix cores; `work` holds the slices and scales (work_bytes long).
// K-batches are split so no int32 sum can overflow, and batches are split so
// one part's slices fit the workspace.
void gemm_int8_emulated(const GemmDesc& d, int slices, void* work, size_t work_bytes, hipStream_t s);

// Workspace one batch of d needs (the least gemm_int8_emulated accepts).
size_t gemm_int8_emulated_workspace(const GemmDesc& d, int slices);

}  // namespace ggpu
