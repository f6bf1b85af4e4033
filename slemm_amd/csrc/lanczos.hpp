// Operator-generic Lanczos seed (the L_Seed recurrences of src/reml.cpp:246-339 with FP64
// scalars and per-group stopping) for callers with their own operator.
#pragma once

#include <functional>
#include <vector>

#include "common.hpp"

namespace slemm {

// out (ldo) = A U (ldu) for the t columns of the seed system, on the given stream.
using LanczosOp = std::function<void(const float* U, int64_t ldu, float* out, int64_t ldo, hipStream_t st)>;
// Called between steps when columns stop (active[k] = 0 for converged groups), so that an
// operator can skip their work; never called while a step is being captured.  Returns true when
// the operator's launches change (another column count): the captured steps are then dropped,
// the next step runs eagerly (so that its buffers exist) and later ones are captured again.
using ActiveFn = std::function<bool(const std::vector<int>& active, hipStream_t st)>;

// Runs the seed systems for the columns of B (n x t, ldb) and returns the seed (owned by the
// caller, released with slemm_gpu_lseed_destroy).  With use_graph, steps after the first replay
// HIP graphs of the operator and Lanczos kernels (op must then launch only kernels on the given
// stream once its buffers exist).  Throws slemm::Error on failure.
slemm_lseed* lseed_run(slemm_geno* h, const LanczosOp& op, const int* col_group, const float* B, int64_t ldb,
                       int t, float tol, int maxit, int keep_cols, int verbose, hipStream_t st, bool use_graph,
                       const ActiveFn& on_active = nullptr);

// Graph replay of Lanczos steps, on unless SLEMM_LANCZOS_GRAPH=0.
bool lanczos_graphs();

}  // namespace slemm
