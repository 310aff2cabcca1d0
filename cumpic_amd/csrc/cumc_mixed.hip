// cumc_mixed.hip -- FP32 instantiations of the reconstruction, Riemann and EMF kernels (the
// mixed-precision form).  Built separately from cumc.hip so that these kernels may contract
// multiply-adds and use the hardware FP32 reciprocal / square root, while the FP64 drop-in
// path keeps every operation separately rounded.  The conserved state, the RK3 combination,
// the conserved-to-primitive recovery and the face-B update stay FP64 (k_tail).
#include "ctx.hpp"

void stage_mixed(cumc_ctx *ctx, int st) { stage<float>(ctx, ctx->F32, st); }

void fluxes_mixed(cumc_ctx *ctx, int dir, double *aos_tmp)
{
    const Lay &L = ctx->L;
    const long ncell = (long)(L.nx + 6) * (L.ny + 6) * (L.nz + 6);
    launch_fluxes<float>(ctx, ctx->F32, dir);
    hipLaunchKernelGGL((k_flux_to_aos<float>), dim3(nblk(ncell, 256)), dim3(256), 0, ctx->stream, L,
                       ctx->F32.F[dir - 1], aos_tmp);
}
