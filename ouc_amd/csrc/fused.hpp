// Fused single-pass engine (one kernel per wavefield per step, batched shots).
#pragma once

#include <vector>

#include "../../include/ouc.h"

struct ouc_ctx;

void fused_init(ouc_ctx *c);
void fused_destroy(ouc_ctx *c);
// One RTM shot through the fused kernels (a batch of one); images to host.
void fused_rtm_one(ouc_ctx *c, const ouc_shot *s, const float *rec_in, float *img_pp, float *img_ps, float *lap_pp,
                   float *lap_ps);
// shot_lap (optional, host [nshots][2][nx*nz]) receives each shot's Laplacian
// PP and PS images.
void fused_rtm_batch(ouc_ctx *c, const ouc_shot *shots, int nshots, float *stack_pp, float *stack_ps,
                     float *per_shot_ms, float *shot_lap = nullptr);
// One job's shots on several contexts of one grid at once (ouc_rtm_concurrent).
void fused_rtm_concurrent(ouc_ctx *const *ctxs, int n, const ouc_shot *shots, int nshots, float *stack_pp,
                          float *stack_ps, float *per_shot_ms);
void fused_rtm_group(ouc_ctx *const *ctxs, int n, const ouc_shot *const *shots, const int32_t *nshots,
                     float *const *stack_pp, float *const *stack_ps);
double fused_own_bytes(const ouc_ctx *c);
// FWI gradient (gradfun with the adsource residual) summed over shots.
void fused_fwi_gradient(ouc_ctx *c, const ouc_shot *shots, int nshots, const float *observed, float *grad_vp,
                        float *grad_vs, float *illum, double *misfit);
// revise_model_end on one array (device memory held for the call only).
void fused_model_update(int device, int64_t n, const float *v, const float *grad, double step, int iter, float *out);
// Forward seismograms [6][nx][nt] of one shot through the fused kernels.
void fused_forward_one(ouc_ctx *c, const ouc_shot *s, float *seis);
// Forward run with V snapshots over a padded-grid rectangle after each listed step.
void fused_snapshots(ouc_ctx *c, const ouc_shot *s, int n, const int32_t *steps, int x0, int z0, int w, int h,
                     float *out);
// RTM with snapshots of the reconstructed source V^t at the listed (decreasing) times.
void fused_recon_snapshots(ouc_ctx *c, const ouc_shot *s, int n, const int32_t *steps, int x0, int z0, int w, int h,
                           float *out);
// Learned-boundary network weights (NULL: the pad without the network).
void fused_set_learned(ouc_ctx *c, const float *weights, int n);
void fused_get_learned_filter(ouc_ctx *c, float *taps);
void fused_set_learned_damping(ouc_ctx *c, float beta);
// Device bytes one shot of a fused batch holds (state, strips, records).
double fused_shot_bytes(const ouc_ctx *c);
// Core tile and its autotuner.
void fused_set_tile(ouc_ctx *c, int cols, int pf);
void fused_get_tile(ouc_ctx *c, int *cols, int *pf);
void fused_autotune(ouc_ctx *c, int nshots, int reps, int *best_cols, int *best_pf, float *ms);
// Forgets the captured time-loop graphs.
void fused_drop_graphs(ouc_ctx *c);
// The CPML context a learned-boundary context falls back to (engine.hip).
ouc_ctx *fallback_ctx(ouc_ctx *c);
