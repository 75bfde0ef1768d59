// Host-side helpers shared by the translation units of libswe_hip.
#pragma once

#include <string>

struct SweContext;

namespace swe {
struct Grid;
// Views of a single-domain context for the other translation units.
const Grid* context_grid(const SweContext* c);
int context_variant(const SweContext* c);
double context_gravity(const SweContext* c);
double context_rate(const SweContext* c, double t);
int context_reduce_blocks(const SweContext* c);
// HIP device the context was created on.
int context_device(const SweContext* c);
// Strip height of the adjoint kernel.
int context_adjoint_rows(const SweContext* c);
// Device scratch of at least n_doubles doubles owned by the context (grown on
// demand, kept until swe_destroy); nullptr if the allocation failed.
double* context_scratch(SweContext* c, size_t n_doubles);

// Resolve the device address of the static scratch of
// swe_apply_residual_correction_noalloc on `device` (which must be current);
// once per device, called by swe_create.
int noalloc_scratch_init(int device);

// Record `msg` as the calling thread's last error (swe_last_error) and return `code`.
int set_error(int code, const std::string& msg);
}  // namespace swe
