// Split-BF16 MP2 kernels (placeholder until the BF16 MFMA path lands).
#include "mp2_bf16.h"

int bf16_prepare(mshqc_hip_ctx* ctx, int, const double*, int, int, const double*, int, int, int,
                 Bf16Operands*) {
    return set_error(ctx, MSHQC_HIP_ERR_UNSUPPORTED, "split-BF16 path not built");
}
int bf16_launch_pair(mshqc_hip_ctx* ctx, int, const PairArgs&, const Bf16Operands&) {
    return set_error(ctx, MSHQC_HIP_ERR_UNSUPPORTED, "split-BF16 path not built");
}
int bf16_gemm(mshqc_hip_ctx* ctx, int, const GemmF64&) {
    return set_error(ctx, MSHQC_HIP_ERR_UNSUPPORTED, "split-BF16 path not built");
}
