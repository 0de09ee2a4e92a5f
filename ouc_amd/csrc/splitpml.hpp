// First-order velocity-stress engine with split PML (forward modelling).
#pragma once

#include <functional>
#include <vector>

#include "../../include/ouc.h"

struct ouc_ctx;

namespace ouc {
// Seismograms [6][nx][nt] (vx, vz, vpx, vpz, vsx, vsz) of one shot; wav is the
// shot's nt-sample source wavelet.
void splitpml_forward(ouc_ctx *c, const ouc_shot *s, const std::vector<float> &wav, float *seis);
// RTM of one shot (EXACT alignment, FP32 strips): the forward run records
// vx, vz ([component][nt][nx]) into rec, `between` may replace and mute them,
// then the backward loop accumulates the imaging sums into the pitched
// planes npp, nps, den (zeroed here).
void splitpml_rtm(ouc_ctx *c, const ouc_shot *s, const std::vector<float> &wav, float *rec,
                  const std::function<void()> &between, float *npp, float *nps, float *den);
} // namespace ouc
