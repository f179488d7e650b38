// LDS-resident Taylor step (placeholder: the streamed path serves every grid).
#pragma once

static int fused_config(salmon_ctx*) { return 0; }
static int fused_set_tables(salmon_ctx*) { return SALMON_OK; }
static void fused_free(salmon_ctx*) {}
template <typename T>
static int fused_step(salmon_ctx* c, const double*, const double*, double, double*) {
  return fail(c, SALMON_ERR_UNSUPPORTED, "fused step not available");
}
