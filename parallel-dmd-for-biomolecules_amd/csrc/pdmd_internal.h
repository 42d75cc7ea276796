// pdmd_internal.h -- library-internal split of pdmd_system_add into its host
// half (validation, topology, the -Drunr H-bond restore, tables, cutoffs;
// no device call) and its device half (uploads, start-up kernel), so that a
// caller inside the library can prepare systems while a launch runs and
// admit them afterwards with device work only.
#pragma once
#include <cstdint>

#include "../../include/pdmd.h"

namespace pdmd {

struct PreparedSystem;

// Host half of pdmd_system_add for ctx's parameter tables; nullptr (and the
// error text of pdmd_last_error) on failure.  The description's arrays are
// read here only.
PreparedSystem* prepared_new(pdmd_ctx* ctx, const pdmd_system_desc* d);
void prepared_free(PreparedSystem* p);
// device bytes pdmd_system_add allocates for it (pdmd_system_bytes_ctx)
int64_t prepared_bytes(const PreparedSystem* p);
// Device half: waits for the context's stream, uploads and starts the
// system up; p is consumed (left empty) whether or not it succeeds.
int prepared_add(pdmd_ctx* ctx, PreparedSystem* p, int32_t* sys_id);

}  // namespace pdmd
