#include "engine.hpp"
#include "fused.hpp"

void fused_init(ouc_ctx *) {}
void fused_destroy(ouc_ctx *) {}
void fused_forward(ouc_ctx *, const ouc_shot *, const std::vector<float> &, bool, int) {
    throw ouc::Error(OUC_UNSUPPORTED, "fused engine not built yet");
}
void fused_backward(ouc_ctx *, const ouc_shot *, const std::vector<float> &) {
    throw ouc::Error(OUC_UNSUPPORTED, "fused engine not built yet");
}
void fused_rtm_batch(ouc_ctx *, const ouc_shot *, int, float *, float *, float *) {
    throw ouc::Error(OUC_UNSUPPORTED, "fused engine not built yet");
}
double fused_own_bytes(const ouc_ctx *) { return 0; }
