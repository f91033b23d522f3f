// The step either side of the projection for a periodic x (the channel, the
// periodic box, the duct), one rank, on device-resident fields:
//   k_rk_update     RK3 stage update of the interior  time_integration.f90:256-258,
//                                                     :394-396, :526-531
//   k_wrap_x/z      periodic ghost planes             boundary_conditions.f90:263-289,
//                                                     :580-602
//   k_robin_y       Robin (slip-length) wall rows     boundary_conditions.f90:784-804
//   k_wrap_y        periodic y ghost rows (y_bc_type = 0, the periodic box)
//                                                     boundary_conditions.f90:762-781
//   k_wall_z        z wall planes (the duct)          boundary_conditions.f90:737-758
// apply_boundary_conditions (:186-260) is the sequence x wrap, z wrap (or
// z walls, :215-223), y walls (or y wrap, :229-233), x wrap, z wrap (none
// with z walls); on one rank its halo exchanges (:170-183) exchange nothing.  Each launch covers the three fields (blockIdx.z); arithmetic is
// kept operation for operation (no contraction), so results equal the CPU
// path bit for bit.
#include <algorithm>

#include "dfp_internal.h"

namespace dfp {

// extents of U (0), V (1), W (2)
struct Field3 {
    double* F[3];
    int n1[3], n2[3], n3[3];
};

static Field3 field3(const Geo& g, double* U, double* V, double* W) {
    Field3 f{};
    f.F[0] = U; f.F[1] = V; f.F[2] = W;
    f.n1[0] = g.nx;  f.n2[0] = g.nyg; f.n3[0] = g.nzg;
    f.n1[1] = g.nxg; f.n2[1] = g.ny;  f.n3[1] = g.nzg;
    f.n1[2] = g.nxg; f.n2[2] = g.nyg; f.n3[2] = g.nz;
    return f;
}

struct RkArgs {
    const double* Fo[3];      // Uo, Vo, Wo
    const double* R[3][3];    // [field][stage term]: interior-shaped right-hand sides
    double c[3];              // rk_coef(stage, 1:stage)
    double dt;
};

// one thread per interior element (i, j) of plane k = blockIdx.y + 2 of field blockIdx.z
template <int STAGE>
__global__ void __launch_bounds__(256) k_rk_update(Field3 f, RkArgs a) {
#pragma clang fp contract(off)
    const int fi = blockIdx.z;
    const int n1 = f.n1[fi], n2 = f.n2[fi], n3 = f.n3[fi];
    const int ni = n1 - 2, nj = n2 - 2;
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int k = 2 + blockIdx.y;
    if (t >= ni * nj || k > n3 - 1) return;
    const int jj = t / ni, ii = t - jj * ni;
    const long e = (ii + 1) + (long)(jj + 1) * n1 + (long)(k - 1) * n1 * n2;  // (i, j, k) = (ii+2, jj+2, k)
    const long r = ii + (long)jj * ni + (long)(k - 2) * ni * nj;             // interior array
    const double fo = a.Fo[fi][e];
    double v;
    if (STAGE == 1) {
        v = fo + a.dt * a.c[0] * a.R[fi][0][r];
    } else if (STAGE == 2) {
        v = fo + a.dt * (a.c[0] * a.R[fi][0][r] + a.c[1] * a.R[fi][1][r]);
    } else {
        v = fo + a.dt * (a.c[0] * a.R[fi][0][r] + a.c[1] * a.R[fi][1][r] + a.c[2] * a.R[fi][2][r]);
    }
    f.F[fi][e] = v;
}

// x ghost planes over (j, k): U faces (1 <- nx-1, nx <- 2), V and W centres
// (1 <- nxg-2, nxg-1 <- 2, nxg <- 3)
__global__ void __launch_bounds__(256) k_wrap_x(Field3 f) {
    const int fi = blockIdx.z;
    const int n1 = f.n1[fi], n2 = f.n2[fi], n3 = f.n3[fi];
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)n2 * n3) return;
    double* row = f.F[fi] + t * n1;  // F(:, j, k)
    // in the reference's statement order, like the z wraps (with nxp >= 2 no
    // copy reads a plane an earlier one wrote)
    if (fi == 0) {
        row[0] = row[n1 - 2];
        row[n1 - 1] = row[1];
    } else {
        row[0] = row[n1 - 3];
        row[n1 - 2] = row[1];
        row[n1 - 1] = row[2];
    }
}

// z ghost planes over (i, j): U and V centres (1 <- nzg-2, nzg-1 <- 2,
// nzg <- 3), W faces (1 <- nz-1, nz <- 2)
__global__ void __launch_bounds__(256) k_wrap_z(Field3 f) {
    const int fi = blockIdx.z;
    const int n1 = f.n1[fi], n2 = f.n2[fi], n3 = f.n3[fi];
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    const long pl = (long)n1 * n2;
    if (t >= pl) return;
    double* F = f.F[fi] + t;  // F(i, j, :) with stride pl
    // in the reference's statement order (with nzp = 1, nzg-1 = 3 is read
    // after it is written)
    if (fi == 2) {
        F[0] = F[(n3 - 2) * pl];
        F[(n3 - 1) * pl] = F[pl];
    } else {
        F[0] = F[(n3 - 3) * pl];
        F[(n3 - 2) * pl] = F[pl];
        F[(n3 - 1) * pl] = F[2 * pl];
    }
}

// y ghost rows over (i, k): V faces (1 <- ny-1, ny <- 2), U and W centres
// (1 <- nyg-2, nyg-1 <- 2, nyg <- 3), in the reference's statement order
__global__ void __launch_bounds__(256) k_wrap_y(Field3 f) {
    const int fi = blockIdx.z;
    const int n1 = f.n1[fi], n2 = f.n2[fi], n3 = f.n3[fi];
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)n1 * n3) return;
    const long k = t / n1, i = t - k * n1;
    double* F = f.F[fi] + i + k * (long)n1 * n2;  // F(i, 1, k), stride n1 in y
    if (fi == 1) {
        F[0] = F[(long)(n2 - 2) * n1];
        F[(long)(n2 - 1) * n1] = F[n1];
    } else {
        F[0] = F[(long)(n2 - 3) * n1];
        F[(long)(n2 - 2) * n1] = F[n1];
        F[(long)(n2 - 1) * n1] = F[2 * (long)n1];
    }
}

// z wall planes over (i, j): W faces no-penetration (1 and nz zero), U and
// V centres no-slip by mirroring (1 <- -2, nzg <- -(nzg-1))
__global__ void __launch_bounds__(256) k_wall_z(Field3 f) {
    const int fi = blockIdx.z;
    const int n1 = f.n1[fi], n2 = f.n2[fi], n3 = f.n3[fi];
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    const long pl = (long)n1 * n2;
    if (t >= pl) return;
    double* F = f.F[fi] + t;  // F(i, j, :) with stride pl
    if (fi == 2) {
        F[0] = 0.0;
        F[(n3 - 1) * pl] = 0.0;
    } else {
        F[0] = -F[pl];
        F[(n3 - 1) * pl] = -F[(n3 - 2) * pl];
    }
}

struct RobinArgs {
    const double* alpha[3];  // alpha_x (nx,2,nzg), alpha_y (nxg,2,nzg), alpha_z (nxg,2,nz)
    double dy1, dy2;         // y(2)-y(1), y(ny)-y(ny-1)
    double dg1, dg2;         // yg(2)-yg(1), yg(nyg)-yg(nyg-1)
};

// wall rows over (i, k): V faces (id 1), U and W centres (id 2)
__global__ void __launch_bounds__(256) k_robin_y(Field3 f, RobinArgs a) {
#pragma clang fp contract(off)
    const int fi = blockIdx.z;
    const int n1 = f.n1[fi], n2 = f.n2[fi], n3 = f.n3[fi];
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)n1 * n3) return;
    const long k = t / n1, i = t - k * n1;
    double* F = f.F[fi] + i + k * (long)n1 * n2;  // F(i, 1, k), stride n1 in y
    const double a1 = a.alpha[fi][i + 2 * k * (long)n1];
    const double a2 = a.alpha[fi][i + n1 + 2 * k * (long)n1];
    // lower wall first, then the upper one, which reads row n2-1 after the
    // lower row is written (they coincide for V with one cell between the walls)
    if (fi == 1) {
        F[0] = a1 * F[n1] / a.dy1 / (a1 / a.dy1 + 1.0);
        F[(long)(n2 - 1) * n1] = a2 * F[(long)(n2 - 2) * n1] / a.dy2 / (a2 / a.dy2 + 1.0);
    } else {
        F[0] = (2.0 * a1 / a.dg1 - 1.0) * F[n1] / (2.0 * a1 / a.dg1 + 1.0);
        F[(long)(n2 - 1) * n1] = (2.0 * a2 / a.dg2 - 1.0) * F[(long)(n2 - 2) * n1] / (2.0 * a2 / a.dg2 + 1.0);
    }
}

hipError_t launch_rk_update(dfp_plan* p, int stage, double* U, double* V, double* W, const double* const Fo[3],
                            const double* const* R[3], const double* c, double dt) {
    const Field3 f = field3(p->geo, U, V, W);
    RkArgs a{};
    for (int fi = 0; fi < 3; ++fi) {
        a.Fo[fi] = Fo[fi];
        for (int s = 0; s < stage; ++s) a.R[fi][s] = R[fi][s];
    }
    for (int s = 0; s < stage; ++s) a.c[s] = c[s];
    a.dt = dt;
    long plane = 0;
    for (int fi = 0; fi < 3; ++fi) plane = std::max(plane, (long)(f.n1[fi] - 2) * (f.n2[fi] - 2));
    int n3 = 0;
    for (int fi = 0; fi < 3; ++fi) n3 = std::max(n3, f.n3[fi] - 2);
    const dim3 grid(ceil_div(plane, 256), n3, 3);
    const int slot = prof_begin(p, "rk_update");
    if (stage == 1) hipLaunchKernelGGL(k_rk_update<1>, grid, dim3(256), 0, p->stream, f, a);
    else if (stage == 2) hipLaunchKernelGGL(k_rk_update<2>, grid, dim3(256), 0, p->stream, f, a);
    else hipLaunchKernelGGL(k_rk_update<3>, grid, dim3(256), 0, p->stream, f, a);
    prof_end(p, slot);
    return hipGetLastError();
}

hipError_t launch_apply_bc(dfp_plan* p, double* U, double* V, double* W, const double* ax, const double* ay,
                           const double* az) {
    const Field3 f = field3(p->geo, U, V, W);
    RobinArgs r{};
    r.alpha[0] = ax; r.alpha[1] = ay; r.alpha[2] = az;
    const auto& y = p->hy;
    const auto& yg = p->hyg;
    r.dy1 = y[1] - y[0];
    r.dy2 = y[y.size() - 1] - y[y.size() - 2];
    r.dg1 = yg[1] - yg[0];
    r.dg2 = yg[yg.size() - 1] - yg[yg.size() - 2];
    long px = 0, pz = 0, py = 0;
    for (int fi = 0; fi < 3; ++fi) {
        px = std::max(px, (long)f.n2[fi] * f.n3[fi]);
        pz = std::max(pz, (long)f.n1[fi] * f.n2[fi]);
        py = std::max(py, (long)f.n1[fi] * f.n3[fi]);
    }
    const int slot = prof_begin(p, "apply_bc");
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_wrap_x, dim3(ceil_div(px, 256), 1, 3), dim3(256), 0, p->stream, f);
        if (p->geo.zbc == 0)
            hipLaunchKernelGGL(k_wrap_z, dim3(ceil_div(pz, 256), 1, 3), dim3(256), 0, p->stream, f);
        else if (pass == 0)
            hipLaunchKernelGGL(k_wall_z, dim3(ceil_div(pz, 256), 1, 3), dim3(256), 0, p->stream, f);
        if (pass == 0 && p->geo.ybc == 0)
            hipLaunchKernelGGL(k_wrap_y, dim3(ceil_div(py, 256), 1, 3), dim3(256), 0, p->stream, f);
        else if (pass == 0)
            hipLaunchKernelGGL(k_robin_y, dim3(ceil_div(py, 256), 1, 3), dim3(256), 0, p->stream, f, r);
    }
    prof_end(p, slot);
    return hipGetLastError();
}

}  // namespace dfp
