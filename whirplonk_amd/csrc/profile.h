// profile.h — WHIR_PROFILE stage rows in the reference's CSV format
// (profiling.hpp:12-142): with WHIR_PROFILE set (not empty and not starting
// with 0, f, F, n or N) the library prints `mode,phase,size,stage,time_ms` rows
// to stderr, once a header, for the same stage vocabulary the reference keeps
// (keep_stage, profiling.hpp:61-77).  Modes: "gpu" for the host-pointer
// drop-ins (the reference's "cuda" rows: witness_gpu_alloc, witness_h2d,
// witness_rs_encode, witness_leaf_hash, witness_d2h, timed with HIP events on
// the call's stream), "prover" for the commit/prove API (witness_encoding,
// merkle_build_total, ood_evaluation, commit_total, prove_total,
// total_prover).  Phases follow the reference's PhaseGuard: "commit" inside
// whir_prover_commit, "prove" inside whir_prover_prove, else "none".
#pragma once
#include <cstddef>

namespace whir_gpu {
namespace prof {

bool enabled();
// one CSV row if enabled and the stage is one the reference keeps
void record(const char* mode, size_t size, const char* stage, double ms);

// the current phase of this thread, restored when the guard ends
struct PhaseGuard {
    explicit PhaseGuard(const char* phase);
    ~PhaseGuard();
    const char* prev;
};

}  // namespace prof
}  // namespace whir_gpu
