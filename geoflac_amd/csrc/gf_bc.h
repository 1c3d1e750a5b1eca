// gf_bc.h -- bc_update (src/bc_update.f90:6-143) on the device: the boundary forces
// the reference accumulates into `force` before fl_node, which then adds the element
// forces to them (nystressbc = 1).
//
//   k_bc_side   two workgroups, one per side wall (nydrsides = 1): the hydrostatic
//               pressure of the wall's element column, a recurrence down the column
//               (src/bc_update.f90:25-55, :63-90).  The per-element density and
//               height terms are formed in parallel; one thread runs the
//               recurrence in the reference's order from LDS.
//   k_bc_force  one thread per node: zero, then the node's share of the side
//               pressure of the elements above and below it, then the applied
//               stresses of its boundary segments (src/bc_update.f90:96-140) in
//               increasing segment order -- the serial loop's order of additions,
//               as a gather (the reference adds with atomics).
#pragma once
#include "gf_kernels.h"

#define BC_MAXNZ 4096   // node rows of a model with hydrostatic side walls (LDS of k_bc_side)

// boundary-node index of node (j,i): left column, right column, then the top and
// bottom rows without their corners; -1 inside
GF_DEV int bc_edge(int j, int i, int nz, int nx)
{
    if (i == 1) return j - 1;
    if (i == nx) return nz + j - 1;
    if (j == 1) return 2 * nz + (i - 2);
    if (j == nz) return 2 * nz + (nx - 2) + (i - 2);
    return -1;
}

template <int M>
__global__ void __launch_bounds__(256) k_bc_side(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    __shared__ double sdpt[BC_MAXNZ], sbeta[BC_MAXNZ];
    const int side = blockIdx.x;  // 0 left, 1 right
    const int ic = side == 0 ? 1 : NX - 1;
    for (int j = 1 + (int)threadIdx.x; j <= NZ - 1; j += blockDim.x) {
        const long e = IE(j, ic);
        const int iph = A.iphase[e];
        const double densT = eff_dens_value<M>(P, ck, A, e, 0);
        const double dh1 = CORD(j, ic, 2) - CORD(j + 1, ic, 2);
        const double dh2 = CORD(j, ic + 1, 2) - CORD(j + 1, ic + 1, 2);
        const double dh = 0.5 * (dh1 + dh2);
        sdpt[j - 1] = densT * P->g * dh;
        sbeta[j - 1] = PHASE(iph).beta;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double rogh = 0.0;
        for (int j = 1; j <= NZ - 1; j++) {
            const double dPT = sdpt[j - 1], b = sbeta[j - 1];
            const double dP = dPT * (1 - b * rogh) / (1 + b / 2 * dPT);
            sdpt[j - 1] = rogh + 0.5 * dP;  // the element's pressure
            rogh = rogh + dP;
        }
    }
    __syncthreads();
    for (int j = 1 + (int)threadIdx.x; j <= NZ - 1; j += blockDim.x)
        A.bc_press[(long)side * (NZ - 1) + (j - 1)] = sdpt[j - 1];
}

template <typename SV>
__global__ void __launch_bounds__(256) k_bc_force(const GfDevParams *__restrict__ P, GfDevClock *ck, GfDevArrays A)
{
    const long t = xcd_lin();
    if (t >= NN) return;
    const int nz = NZ, nx = NX;
    const int j = (int)((unsigned)t % (unsigned)nz) + 1, i = (int)((unsigned)t / (unsigned)nz) + 1;
    double f1 = 0.0, f2 = 0.0;
    const int ed = bc_edge(j, i, nz, nx);
    if (ed >= 0) {
        if (P->nydrsides == 1 && (i == 1 || i == nx)) {
            const int side = i == 1 ? 0 : 1;
            // elements j-1 (this node is its lower node) and j (its upper node)
            for (int je = j - 1; je <= j; je++) {
                if (je < 1 || je > nz - 1) continue;
                const double press_norm = -A.bc_press[(long)side * (nz - 1) + (je - 1)];
                const double dlx = CORD(je + 1, i, 1) - CORD(je, i, 1);
                const double dly = CORD(je + 1, i, 2) - CORD(je, i, 2);
                if (side == 0) {
                    f1 = f1 + 0.5 * press_norm * dly;
                    f2 = f2 - 0.5 * press_norm * dlx;
                } else {
                    f1 = f1 - 0.5 * press_norm * dly;
                    f2 = f2 + 0.5 * press_norm * dlx;
                }
            }
        }
        for (int q = A.bc_estart[ed]; q < A.bc_estart[ed + 1]; q++) {
            const int sg = A.bc_elist[q] >> 1;
            const long na = A.bc_nodes[2 * sg], nb = A.bc_nodes[2 * sg + 1];
            const double dlx = A.cord[nb] - A.cord[na];
            const double dly = A.cord[nb + NN] - A.cord[na + NN];
            const double sn = A.bc_str[2 * sg], ss = A.bc_str[2 * sg + 1];
            f1 = f1 - 0.5 * sn * dly;  // normal component
            f2 = f2 + 0.5 * sn * dlx;
            f1 = f1 + 0.5 * ss * dlx;  // shear component
            f2 = f2 + 0.5 * ss * dly;
        }
    }
    SV *force = sv_ptr<SV>(A.force, A.forcef);
    force[t] = SV(f1);
    force[t + NN] = SV(f2);
}
