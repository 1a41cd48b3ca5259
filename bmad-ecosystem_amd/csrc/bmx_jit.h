// bmx_jit.h -- runtime-compiled cell kernels (bmx_jit.cpp), host interface.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "bmx_internal.h"

namespace bmx {

// The lattice's repeating cell: `len` ops whose types and flags equal those
// at prog[rep ..]; n_cells non-overlapping runs of it are flagged F_CELL.
struct CellPlan {
  int len = 0;
  size_t rep = 0;
  size_t n_cells = 0;
  std::string key;                // the cell's op types and flags, the compile cache key
  std::vector<int32_t> cell_prv;  // per op: the last cell start at or before it (-1: none)
};

struct JitKernels {
  hipFunction_t bunch, multi, grid;
};

// Finds the cell and flags its runs in prog (clears F_CELL everywhere first).
// False when the program has no cell covering at least half of its ops, or
// BMX_JIT=0.
bool plan_cells(std::vector<DevOp>& prog, CellPlan* plan);

// The cell kernels for this plan on `device` (compiled on first use, then
// cached for the process); nullptr and *err on failure.
const JitKernels* jit_kernels(int device, const std::vector<DevOp>& prog, const CellPlan& plan, std::string* err);

}  // namespace bmx
