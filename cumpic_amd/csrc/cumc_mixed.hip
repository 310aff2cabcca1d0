// cumc_mixed.hip -- FP32 instantiations of the reconstruction, Riemann and EMF kernels (the
// mixed-precision form).  Built separately from cumc.hip so that these kernels may contract
// multiply-adds and use the hardware FP32 reciprocal / square root, while the FP64 drop-in
// path keeps every operation separately rounded.  The conserved state, the RK3 combination,
// the conserved-to-primitive recovery and the face-B update stay FP64 (k_tail).
#include "ctx.hpp"

void stage_mixed(cumc_ctx *ctx, int st, bool comb) { stage<float>(ctx, ctx->F32, st, comb); }

// the reference's ghost passes (and torus inflow clamps) on the FP32 primitive planes
void launch_ghost32(cumc_ctx *ctx)
{
    const Lay &L = ctx->L;
    hipStream_t s = ctx->stream;
    const Fields<float> &G = ctx->F32;
    const long nx6 = L.nx + 6, ny6 = L.ny + 6;
    hipLaunchKernelGGL(k_ghost32<0>, dim3(nblk(ghost_threads<0>(L), 256)), dim3(256), 0, s, L, G.P32, G.EPS32, G.BC32,
                       ctx->M);
    hipLaunchKernelGGL(k_ghost32<1>, dim3(nblk(ghost_threads<1>(L), 256)), dim3(256), 0, s, L, G.P32, G.EPS32, G.BC32,
                       ctx->M);
    hipLaunchKernelGGL(k_ghost32<2>, dim3(nblk(ghost_threads<2>(L), 256)), dim3(256), 0, s, L, G.P32, G.EPS32, G.BC32,
                       ctx->M);
    if (ctx->M.model == CUMC_MODEL_TORUS_CYN || ctx->M.model == CUMC_MODEL_TORUS_SPH) {
        const long nyz = ny6 * (L.nz + 6) * 3, nxy = nx6 * ny6 * 3;
        hipLaunchKernelGGL(k_custom_boundary_x<float>, dim3(nblk(nyz, 256)), dim3(256), 0, s, L, G.P32);
        if (ctx->M.model == CUMC_MODEL_TORUS_CYN)
            hipLaunchKernelGGL(k_custom_boundary_z<float>, dim3(nblk(nxy, 256)), dim3(256), 0, s, L, G.P32);
    }
}

void fluxes_mixed(cumc_ctx *ctx, int dir, double *aos_tmp)
{
    const Lay &L = ctx->L;
    const long ncell = (long)(L.nx + 6) * (L.ny + 6) * (L.nz + 6);
    launch_fluxes<float>(ctx, ctx->F32, dir);
    hipLaunchKernelGGL((k_flux_to_aos<float>), dim3(nblk(ncell, 256)), dim3(256), 0, ctx->stream, L,
                       ctx->F32.F[dir - 1], aos_tmp);
}
