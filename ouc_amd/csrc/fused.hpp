// Fused single-pass engine (one kernel per wavefield per step, batched shots).
#pragma once

#include <vector>

#include "../../include/ouc.h"

struct ouc_ctx;

void fused_init(ouc_ctx *c);
void fused_destroy(ouc_ctx *c);
// One RTM shot through the fused kernels (a batch of one); images to host.
void fused_rtm_one(ouc_ctx *c, const ouc_shot *s, float *img_pp, float *img_ps, float *lap_pp, float *lap_ps);
void fused_rtm_batch(ouc_ctx *c, const ouc_shot *shots, int nshots, float *stack_pp, float *stack_ps,
                     float *per_shot_ms);
double fused_own_bytes(const ouc_ctx *c);
