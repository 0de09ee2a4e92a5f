// First-order velocity-stress engine with split PML (forward modelling).
#pragma once

#include <vector>

#include "../../include/ouc.h"

struct ouc_ctx;

namespace ouc {
// Seismograms [6][nx][nt] (vx, vz, vpx, vpz, vsx, vsz) of one shot; wav is the
// shot's nt-sample source wavelet.
void splitpml_forward(ouc_ctx *c, const ouc_shot *s, const std::vector<float> &wav, float *seis);
} // namespace ouc
