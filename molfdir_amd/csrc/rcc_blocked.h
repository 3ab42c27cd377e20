// The symmetry-blocked CCSD engine (rcc_blocked.hip) as the solver sees it.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include "rcc_blocked_plan.h"

namespace rcc {

struct BlkEngine;
// Plan, device arrays and index maps for one symmetry (current device).
// Returns 1 when the symmetry has no blocked form (irreps that are not their
// own inverse), -1 on failure.
int blk_create(const blk::SymIn& s, BlkEngine** out, std::string& err);
void blk_destroy(BlkEngine* E);
// Arrays held outside the engine (the VVVV and VOVV blocks of the handle);
// the product descriptors are built once both are bound.
int blk_bind(BlkEngine* E, int id, double2* p, std::string& err);
double2* blk_ptr(BlkEngine* E, int id);
int64_t blk_size(const BlkEngine* E, int id);
int64_t blk_bytes(const BlkEngine* E);
// Denominators and frozen masks from the orbital energies (host array, C1 order).
int blk_set_eps(BlkEngine* E, const double* eps, std::string& err);
// T changed: tau, the ring form and X are recomputed on demand.
void blk_invalidate(BlkEngine* E);
// Runs one program (blk::Prog), computing tau / ring / X first where the term
// needs them; enqueues on st, no host synchronisation.
int blk_run(BlkEngine* E, int prog, int gauss, hipStream_t st, std::string& err);
// Parameters of tau(S) (RK_MAPBIL) for the batched CCENG.
void blk_tau_s_params(const BlkEngine* E, void* q);
// Real flops (classical four-product count) of one iteration's products.
double blk_iteration_flops(const BlkEngine* E);

}  // namespace rcc
