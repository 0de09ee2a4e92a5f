// Fused single-pass engine (one kernel per wavefield per step, batched shots).
#pragma once

#include <vector>

#include "../../include/ouc.h"

struct ouc_ctx;

void fused_init(ouc_ctx *c);
void fused_destroy(ouc_ctx *c);
void fused_forward(ouc_ctx *c, const ouc_shot *s, const std::vector<float> &wav, bool rtm, int ncomp);
void fused_backward(ouc_ctx *c, const ouc_shot *s, const std::vector<float> &wav);
void fused_rtm_batch(ouc_ctx *c, const ouc_shot *shots, int nshots, float *stack_pp, float *stack_ps,
                     float *per_shot_ms);
double fused_own_bytes(const ouc_ctx *c);
