// cpu_step_harness.cpp -- runs the device step arithmetic (rtm_step.h) on the
// CPU so the per-point fused update can be checked against the oracle
// without a GPU.  Built by tests/test_cpu_harness.py with
// g++ -O2 -ffp-contract=off; test infrastructure only.
#include <cstddef>

#include "rtm_step.h"

template <typename R>
static void run(const int *g, const R *hs, const R *u, R *U, const R *Cf, const R *ax, const R *bx,
                const R *az, const R *bz, const R *pXi, R *pXo, R *cX, const R *pZi, R *pZo, R *cZ,
                const int *recmap, const float *seis_row, R *snap, R *I, R *Is, R *Ic, R *Isc)
{
    rtm::StepArgs<R> A{};
    A.X = g[0]; A.Z = g[1]; A.P = g[2];
    A.TL = g[3]; A.TR = g[4]; A.TU = g[5]; A.TD = g[6];
    A.nx = g[7]; A.nz = g[8]; A.mode = g[9]; A.src_j = g[10]; A.rec_row = g[11]; A.nrec = g[12];
    A.h1 = hs[0]; A.h2 = hs[1]; A.h3 = hs[2]; A.src_w = hs[3];
    A.u = u; A.U = U; A.C = Cf;
    A.ax = ax; A.bx = bx; A.az = az; A.bz = bz;
    A.psiXi = pXi; A.psiXo = pXo; A.csiX = cX;
    A.psiZi = pZi; A.psiZo = pZo; A.csiZ = cZ;
    A.recmap = recmap; A.seis_row = seis_row;
    A.snap = snap; A.I = I; A.Is = Is; A.Ic = Ic; A.Isc = Isc;
    for (int k = rtm::kHalf; k < A.Z - rtm::kHalf; k++)
        for (int i = rtm::kHalf; i < A.X - rtm::kHalf; i++) rtm::step_point(A, i, k);
    if (g[13])   // outer-column top-strip pass, as k_top_edges
        for (int k = rtm::kHalf; k < A.TU - rtm::kHalf; k++)
            for (int c = 0; c < 8; c++) rtm::cpml_top_edge_point(A, c < 4 ? c : A.X - 8 + c, k);
}

extern "C" {
void hx_step_f32(const int *g, const float *hs, const float *u, float *U, const float *Cf,
                 const float *ax, const float *bx, const float *az, const float *bz, const float *pXi,
                 float *pXo, float *cX, const float *pZi, float *pZo, float *cZ, const int *recmap,
                 const float *seis_row, float *snap, float *I, float *Is, float *Ic, float *Isc)
{
    run<float>(g, hs, u, U, Cf, ax, bx, az, bz, pXi, pXo, cX, pZi, pZo, cZ, recmap, seis_row, snap, I,
               Is, Ic, Isc);
}

void hx_step_f64(const int *g, const double *hs, const double *u, double *U, const double *Cf,
                 const double *ax, const double *bx, const double *az, const double *bz,
                 const double *pXi, double *pXo, double *cX, const double *pZi, double *pZo,
                 double *cZ, const int *recmap, const float *seis_row, double *snap, double *I,
                 double *Is, double *Ic, double *Isc)
{
    run<double>(g, hs, u, U, Cf, ax, bx, az, bz, pXi, pXo, cX, pZi, pZo, cZ, recmap, seis_row, snap, I,
                Is, Ic, Isc);
}
}
