// bmx_tangent.h -- exact transfer matrices (bmx_tangent.hip), host interface.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bmx_internal.h"

namespace bmx {

// Queues the transfer-matrix kernel on `stream`: n particles (d_in, n x 6)
// through [lo1, hi1) then [lo2, hi2) of the op program for n_turns turns;
// d_out n x 6 final coordinates, d_mat n x 36 row-major Jacobians
// d out_r / d in_q (zeros for a lost particle), d_state the final states.
hipError_t launch_transfer(hipStream_t stream, const DevOp* ops, int lo1, int hi1, int lo2, int hi2,
                           const LatticeConsts& lc, int n_turns, const double* d_in, double* d_out, double* d_mat,
                           int32_t* d_state, int64_t n);

}  // namespace bmx
