// internal.h -- host-side helpers shared by the pipeline and the request server.
#pragma once
#include <cstddef>

#include "../../include/uncertainty_amd.h"

size_t unc_ctx_held_bytes(unc_context *ctx);  // device bytes the context keeps allocated
size_t unc_ctx_ws_bytes(unc_context *ctx);
size_t unc_ctx_io_bytes(unc_context *ctx);
void unc_ctx_release(unc_context *ctx);       // free the context's cached buffers
void unc_ctx_set_fork(unc_context *ctx, bool on);  // Z^+ recursion on a second stream or not
size_t unc_io_bytes_for(const unc_options *o, int num_rows, size_t nnz);
