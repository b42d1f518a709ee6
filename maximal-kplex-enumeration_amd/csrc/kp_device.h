// kp_device.h -- host-callable entry points of the HIP part of libkplex.
#pragma once
#include "kp_internal.h"

struct kp_prepared;

namespace kp {

int  dev_prepare(const kp_graph *g, const kp_params *p, kp_prepared **out);
int  dev_run(kp_prepared *h, kp_result *r);
int  dev_enumerate(const kp_graph *g, const kp_params *p, kp_result *r);
int  dev_enumerate_noalloc(const kp_graph *g, const kp_params *p, kp_result *r);
int  dev_batch(const kp_graph *const *g, const kp_params *p, size_t n, kp_result *r);
int  dev_autotune(kp_prepared *h, kp_tune *out);
int  dev_root_subgraph(kp_prepared *h, int64_t pos, uint32_t *block, uint32_t *n, uint32_t *hop, uint32_t *left,
                       uint32_t *l, uint32_t left_cap, uint8_t *pairs, size_t pairs_cap);
void dev_set_tune(const kp_tune *t);
void dev_free(kp_prepared *h);
int  dev_set_window(kp_prepared *h, int64_t root_begin, int64_t root_end);
uint64_t dev_bytes(const kp_prepared *h);
int  dev_info(char *buf, size_t len);
const char *last_error();

} // namespace kp
