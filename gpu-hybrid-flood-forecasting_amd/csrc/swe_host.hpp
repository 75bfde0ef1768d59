// Host-side helpers shared by the translation units of libswe_hip.
#pragma once

#include <cstddef>
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

// Bytes of the device buffer of prepare_static_record's Manning classes, and
// where in it the class table (256 floats, n^2 of class c) lies.
size_t class_buf_bytes();
const float* class_values(const void* cls_mem);
// Loop start of a packed fast loop on grid G, from its current arrays, queued
// on `stream`: sm[0] = FP64 sum of the rain multiplier over active cells,
// sm[1] = 1 if it takes one bit pattern there, sm[2] = that value; with
// cls_mem (class_buf_bytes() of device memory) also sm[3] = the static record
// (kStaticCls if allow_classes and at most 255 Manning values, else kStatic8
// or kStatic12), written to G.pk, with the class table in G.cls_tab's place
// in cls_mem.  sum_part and bits_part: device scratch of max_blocks doubles
// each.
int prepare_static_record(const Grid& G, void* cls_mem, double* sm, int allow_classes,
                          double* sum_part, unsigned* bits_part, int max_blocks, void* stream);

// Record `msg` as the calling thread's last error (swe_last_error) and return `code`.
int set_error(int code, const std::string& msg);
}  // namespace swe
