// Host-side helpers shared by the translation units of libswe_hip.
#pragma once

#include <string>

struct SweContext;

namespace swe {
struct Grid;
// Read-only views of a single-domain context for the adjoint's launcher.
const Grid* context_grid(const SweContext* c);
int context_variant(const SweContext* c);
double context_gravity(const SweContext* c);
double context_rate(const SweContext* c, double t);

// Record `msg` as the calling thread's last error (swe_last_error) and return `code`.
int set_error(int code, const std::string& msg);
}  // namespace swe
