// flow.hip — least-squares gradient, field correction and a device-resident
// SIMPLE / PISO outer iteration.
//
// Reference: LeastSquaresGradient (src/solver/Gradient.cpp:42-136),
// SimpleAlgorithm::correct_fields / solve_pressure_correction_and_correct /
// iterate / advance_time_step (src/solver/SimpleAlgorithm.cpp:450-519,
// 521-575) and PisoAlgorithm::iterate (src/solver/PisoAlgorithm.cpp:27-70).
//
// The gradient and the correction are cell-centric gathers in the style of
// the assembly kernels: a thread owns a cell and walks its faces in ascending
// face id, the order in which the reference's face loops reach that cell, so
// with contraction off every sum is the reference's sum bit for bit.  The
// outer iteration keeps U, p, the face flux, the momentum diagonal and both
// matrices in HBM; the host only sequences launches and the two solvers.
#include "checkpoint.hpp"
#include "mesh.hpp"
#include "solver.hpp"
#include "stream.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace mi
{

#pragma clang fp contract(off)

namespace
{

// Inverse of the symmetric 3x3 moment matrix (xx xy xz yy yz zz) by its
// adjugate, zero when |det| < 1e-30 (Gradient.cpp:16-33).
__device__ __forceinline__ void inv_sym3(const double (&a)[6], double* out)
{
    const double xx = a[0], xy = a[1], xz = a[2], yy = a[3], yz = a[4], zz = a[5];
    const double det =
        xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    if (fabs(det) < 1e-30) {
        for (int k = 0; k < 6; ++k)
            out[k] = 0.0;
        return;
    }
    const double inv = 1.0 / det;
    out[0] = (yy * zz - yz * yz) * inv;
    out[1] = -(xy * zz - yz * xz) * inv;
    out[2] = (xy * yz - yy * xz) * inv;
    out[3] = (xx * zz - xz * xz) * inv;
    out[4] = -(xx * yz - xy * xz) * inv;
    out[5] = (xx * yy - xy * xy) * inv;
}

__device__ __forceinline__ void ls_moment(double dx, double dy, double dz, double (&a)[6])
{
    const double w2 = 1.0 / (dx * dx + dy * dy + dz * dz + 1e-30);
    a[0] += w2 * dx * dx;
    a[1] += w2 * dx * dy;
    a[2] += w2 * dx * dz;
    a[3] += w2 * dy * dy;
    a[4] += w2 * dy * dz;
    a[5] += w2 * dz * dz;
}

// Moment matrices (Gradient.cpp:42-91): interior faces add the centroid
// offset to the owner and its negative to the neighbour; boundary faces add
// the face-centroid offset to the owner.
__global__ __launch_bounds__(kBlock) void k_ls_setup(MeshDev m, double* __restrict__ minv)
{
    const int c = int(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= m.n_cells)
        return;
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
        const int f = m.cf_face[k];
        const int o = m.owner[f];
        const int nb = m.nbr[f];
        if (nb < 0) {
            ls_moment(m.fcx[f] - m.ccx[o], m.fcy[f] - m.ccy[o], m.fcz[f] - m.ccz[o], a);
        } else {
            const double dx = m.ccx[nb] - m.ccx[o];
            const double dy = m.ccy[nb] - m.ccy[o];
            const double dz = m.ccz[nb] - m.ccz[o];
            if (c == o)
                ls_moment(dx, dy, dz, a);
            else
                ls_moment(-dx, -dy, -dz, a);
        }
    }
    inv_sym3(a, minv + size_t(c) * 6);
}

// grad(phi) of cell c (Gradient.cpp:93-136).  Boundary faces add
// w2 * d * 0.0, a signed zero, to sums that start at +0.0: no change, so they
// are skipped.
__device__ __forceinline__ void ls_grad_cell(const MeshDev& m, const double* __restrict__ minv,
                                             const double* __restrict__ phi, int c, double& gx,
                                             double& gy, double& gz)
{
    double bx = 0.0, by = 0.0, bz = 0.0;
    for (int k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
        const int f = m.cf_face[k];
        const int nb = m.nbr[f];
        if (nb < 0)
            continue;
        const int o = m.owner[f];
        const double dx = m.ccx[nb] - m.ccx[o];
        const double dy = m.ccy[nb] - m.ccy[o];
        const double dz = m.ccz[nb] - m.ccz[o];
        const double w2 = 1.0 / (dx * dx + dy * dy + dz * dz + 1e-30);
        const double dphi = phi[nb] - phi[o];
        const double d = (c == o) ? dphi : -dphi;
        bx += w2 * dx * d;
        by += w2 * dy * d;
        bz += w2 * dz * d;
    }
    const double* mi = minv + size_t(c) * 6;
    const double m0 = mi[0], m1 = mi[1], m2 = mi[2], m3 = mi[3], m4 = mi[4], m5 = mi[5];
    gx = m0 * bx + m1 * by + m2 * bz;
    gy = m1 * bx + m3 * by + m4 * bz;
    gz = m2 * bx + m4 * by + m5 * bz;
}

__global__ __launch_bounds__(kBlock) void k_ls_gradient(MeshDev m, const double* __restrict__ minv,
                                                        const double* __restrict__ phi,
                                                        double* __restrict__ gx,
                                                        double* __restrict__ gy,
                                                        double* __restrict__ gz)
{
    const int c = int(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= m.n_cells)
        return;
    double x, y, z;
    ls_grad_cell(m, minv, phi, c, x, y, z);
    gx[c] = x;
    gy[c] = y;
    gz[c] = z;
}

// Momentum pressure source of component comp: out = grad(p)_comp * V
// (SimpleAlgorithm.cpp:307-315 subtracts it from the rhs).
__global__ __launch_bounds__(kBlock) void k_ls_source(MeshDev m, const double* __restrict__ minv,
                                                      const double* __restrict__ p,
                                                      double* __restrict__ sx,
                                                      double* __restrict__ sy,
                                                      double* __restrict__ sz)
{
    const int c = int(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= m.n_cells)
        return;
    double x, y, z;
    ls_grad_cell(m, minv, p, c, x, y, z);
    const double v = m.vol[c];
    sx[c] = x * v;
    sy[c] = y * v;
    sz[c] = z * v;
}

// Cell part of correct_fields (SimpleAlgorithm.cpp:457-472):
// p += urfP p' ; U -= V / max(aP, 1e-30) grad(p').
__global__ __launch_bounds__(kBlock) void k_correct_cells(MeshDev m,
                                                          const double* __restrict__ minv,
                                                          double urfP,
                                                          const double* __restrict__ aP,
                                                          const double* __restrict__ pp,
                                                          double* __restrict__ p,
                                                          double* __restrict__ Ux,
                                                          double* __restrict__ Uy,
                                                          double* __restrict__ Uz)
{
    const int c = int(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= m.n_cells)
        return;
    p[c] += urfP * pp[c];
    double gx, gy, gz;
    ls_grad_cell(m, minv, pp, c, gx, gy, gz);
    const double w = m.vol[c] / fmax(aP[c], 1e-30);
    Ux[c] -= w * gx;
    Uy[c] -= w * gy;
    Uz[c] -= w * gz;
}

// Face part of correct_fields (SimpleAlgorithm.cpp:474-488).
__global__ __launch_bounds__(kBlock) void k_correct_faces(MeshDev m, double rho,
                                                          const double* __restrict__ aP,
                                                          const double* __restrict__ pp,
                                                          double* __restrict__ phi)
{
    const int f = int(blockIdx.x) * kBlock + threadIdx.x;
    if (f >= m.n_faces)
        return;
    const int o = m.owner[f], nb = m.nbr[f];
    if (nb < 0)
        return;
    const double Ax = m.ax[f], Ay = m.ay[f], Az = m.az[f];
    const double Amag2 = Ax * Ax + Ay * Ay + Az * Az;
    const double dx = m.ccx[nb] - m.ccx[o];
    const double dy = m.ccy[nb] - m.ccy[o];
    const double dz = m.ccz[nb] - m.ccz[o];
    const double ddotA = fabs(dx * Ax + dy * Ay + dz * Az);
    const double aPf = 0.5 * (aP[o] + aP[nb]);
    const double coef = rho * Amag2 / fmax(aPf * ddotA, 1e-30);
    phi[f] += coef * (pp[o] - pp[nb]);
}

// aP[c] = a_cc (the diagonal captured after component 0: SimpleAlgorithm.cpp:349-353)
__global__ __launch_bounds__(kBlock) void k_take_diag(int n, const int* __restrict__ diag,
                                                      const double* __restrict__ val,
                                                      double* __restrict__ aP)
{
    const int c = int(blockIdx.x) * kBlock + threadIdx.x;
    if (c < n)
        aP[c] = val[diag[c]];
}

// partials of ||A x - b||^2 (solve_momentum_component: SimpleAlgorithm.cpp:491-507)
struct EpiResid
{
    const double* b;
    double bi = 0.0, acc = 0.0;
    __device__ __forceinline__ void pre(int i) { bi = b[i]; }
    __device__ __forceinline__ void row(int, double s)
    {
        const double d = s - bi;
        acc += d * d;
    }
};

__global__ __launch_bounds__(kBlock) void k_resid(StreamMat<double> A, const double* __restrict__ x,
                                                  const double* __restrict__ b, double* partials)
{
    EpiResid e{b};
    stream_rows(A, x, e);
    double v[1] = {e.acc};
    block_partials<1>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_sumsq(int n, const double* __restrict__ v,
                                                  double* partials)
{
    const RowRange rr = block_rows(n, gridDim.x, 1);
    double a[1] = {0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock)
        a[0] += v[i] * v[i];
    block_partials<1>(a, partials);
}

__global__ __launch_bounds__(kBlock) void k_fill(int n, double v, double* __restrict__ x)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n)
        x[i] = v;
}

} // namespace

#pragma clang fp contract(on)

// Builds the mesh's moment matrices once.
int ensure_ls(mi_mesh* M, hipStream_t s)
{
    if (M->ls_minv.p || M->n_cells == 0)
        return MI_OK;
    MI_TRY(M->ls_minv.alloc(size_t(M->n_cells) * 6));
    k_ls_setup<<<grid_for(M->n_cells, kBlock), kBlock, 0, s>>>(M->dev(), M->ls_minv.p);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

int launch_ls_gradient(mi_mesh* M, const double* phi, double* gx, double* gy, double* gz,
                       hipStream_t s)
{
    MI_TRY(ensure_ls(M, s));
    if (M->n_cells > 0) {
        k_ls_gradient<<<grid_for(M->n_cells, kBlock), kBlock, 0, s>>>(M->dev(), M->ls_minv.p,
                                                                      phi, gx, gy, gz);
        MI_HIP(hipGetLastError());
    }
    return MI_OK;
}

int launch_correct_fields(mi_mesh* M, double urfP, double rho, const double* aP, const double* pp,
                          double* p, double* Ux, double* Uy, double* Uz, double* phi,
                          hipStream_t s)
{
    MI_TRY(ensure_ls(M, s));
    if (M->n_cells > 0) {
        k_correct_cells<<<grid_for(M->n_cells, kBlock), kBlock, 0, s>>>(
            M->dev(), M->ls_minv.p, urfP, aP, pp, p, Ux, Uy, Uz);
        MI_HIP(hipGetLastError());
    }
    if (M->n_faces > 0 && phi) {
        k_correct_faces<<<grid_for(M->n_faces, kBlock), kBlock, 0, s>>>(M->dev(), rho, aP, pp,
                                                                        phi);
        MI_HIP(hipGetLastError());
    }
    return MI_OK;
}

} // namespace mi

using namespace mi;

// Device-resident state of SimpleAlgorithm / PisoAlgorithm.
struct mi_flow
{
    mi_mesh* M = nullptr; // borrowed
    mi_flow_cfg cfg{};
    hipStream_t s = nullptr;
    int n = 0, nf = 0;
    int time_step = 0;
    DevBuf<double> U[3], Un[3], Unm1[3], p, phi, aP, pp, rhs, src[3], dt, bc, partials, scal;
    DevBuf<double> den; // SIMPLEC denominators
    double* h_scal = nullptr; // pinned
    std::unique_ptr<mi_csr> Am, Ap;
    std::unique_ptr<mi_solver> sm, sp;
    mi_batch* mb = nullptr;   // momentum_batch: the three components' batched solve
    DevBuf<double> rhs3, x3;  // their right-hand sides and solutions, component-major
    ~mi_flow()
    {
        if (s)
            (void)hipStreamSynchronize(s);
        if (mb)
            (void)mi_batch_destroy(mb);
        sm.reset();
        sp.reset();
        if (h_scal)
            (void)hipHostFree(h_scal);
        if (s)
            (void)hipStreamDestroy(s);
    }
};

namespace
{

int flow_norm(mi_flow* F, const double* v, double* out_dev)
{
    const unsigned G = range_grid(F->n, kBlock);
    k_sumsq<<<G, kBlock, 0, F->s>>>(F->n, v, F->partials.p);
    MI_HIP(hipGetLastError());
    return launch_fin_sum(F->partials.p, G, true, out_dev, F->s);
}

int flow_resid(mi_flow* F, const mi_csr* A, const double* x, const double* b, double* out_dev)
{
    const Pattern& P = *A->pat;
    StreamMat<double> S{int(P.n), P.rb.nrb, P.rb.blk.p, P.rp.p, P.ci.p, A->val.p, nullptr, P.rb.lpr};
    S.hdr = P.rb.hdr.p;
    const unsigned G = stream_grid(P.rb.nrb);
    k_resid<<<G, kBlock, 0, F->s>>>(S, x, b, F->partials.p);
    MI_HIP(hipGetLastError());
    return launch_fin_sum(F->partials.p, G, true, out_dev, F->s);
}

// Device solve of A x = b with x0 = 0 (the CPU solvers' start, LinearSolvers.cpp:424):
// the solver's stream waits for the flow's work by a host synchronisation.
int flow_solve(mi_flow* F, mi_solver* S, const double* b, double* x)
{
    MI_HIP(hipStreamSynchronize(F->s));
    mi_stats st{};
    const int rc = mi_solver_solve_device(S, b, x, &st);
    if (rc != MI_OK && rc != MI_ERR_NOT_CONVERGED && rc != MI_ERR_BREAKDOWN)
        return rc;
    return MI_OK; // the reference's outer loop goes on whatever the solver returned
}

// The momentum predictor's three solves as one batched solve: the components
// share the matrix (only the rhs depends on the component), so the three
// assemblies write the same values; each rhs goes to its third of rhs3, the
// batch solves the three systems together from x = 0 with BiCGStab and one
// ILU(0) factorization of the shared matrix (the matrix and its factors are
// streamed once per launch for the three right-hand sides), and each
// residual ||A U_k - b_k|| is formed as in the one-at-a-time path.  BiCGStab
// whatever momentum.kind: on the 96^3 momentum system the batched
// BiCGStab-ILU(0) solve of the three right-hand sides takes 6.1 ms (31
// iterations), the batched GMRES(30)-ILU(0) one 18.7 ms (58 Arnoldi steps).
int momentum_batched(mi_flow* F, mi_flow_residuals* res, bool transient, bool bdf2)
{
    mi_flow_cfg& c = F->cfg;
    mi_mesh* M = F->M;
    const size_t n = size_t(F->n);
    for (int k = 0; k < 3; ++k)
        MI_TRY(launch_assemble_momentum(M, k, c.mu, c.rho, F->phi.p, F->src[k].p,
                                        transient ? F->dt.p : nullptr, F->Un[k].p,
                                        bdf2 ? F->Unm1[k].p : nullptr, F->U[k].p, c.urfU, F->bc.p,
                                        F->Am->val.p, F->rhs3.p + size_t(k) * n, F->s));
    F->Am->values_changed(false);
    k_take_diag<<<grid_for(F->n, kBlock), kBlock, 0, F->s>>>(F->n, M->pat->diag.p, F->Am->val.p,
                                                             F->aP.p);
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(F->s));
    if (!F->mb)
        MI_TRY(mi_batch_create_ex(F->Am.get(), 3, MI_BATCH_SHARED_VALUES | MI_BATCH_ILU0, &F->mb));
    MI_TRY(mi_batch_set_matrix(F->mb, F->Am.get()));
    mi_solver_cfg bc = c.momentum;
    bc.kind = MI_SOLVER_BICGSTAB;
    bc.preconditioner = MI_PC_ILU;
    int its[3];
    double rel[3];
    const int rc = mi_batch_solve_device(F->mb, &bc, F->rhs3.p, F->x3.p, its, rel);
    if (rc != MI_OK && rc != MI_ERR_NOT_CONVERGED)
        return rc; // the reference's outer loop goes on whatever the solver returned
    for (int k = 0; k < 3; ++k) {
        MI_HIP(hipMemcpyAsync(F->U[k].p, F->x3.p + size_t(k) * n, n * sizeof(double),
                              hipMemcpyDeviceToDevice, F->s));
        MI_TRY(flow_resid(F, F->Am.get(), F->U[k].p, F->rhs3.p + size_t(k) * n, F->scal.p + k));
    }
    MI_HIP(hipMemcpyAsync(F->h_scal, F->scal.p, 3 * sizeof(double), hipMemcpyDeviceToHost, F->s));
    MI_HIP(hipStreamSynchronize(F->s));
    for (int k = 0; k < 3; ++k)
        res->mom[k] = F->h_scal[k];
    return MI_OK;
}

// Momentum predictor: face fluxes, then per component assemble, solve and the
// residual ||A U_k - b|| (SimpleAlgorithm.cpp:527-531 and 491-507).
int momentum_predictor(mi_flow* F, mi_flow_residuals* res)
{
    mi_flow_cfg& c = F->cfg;
    mi_mesh* M = F->M;
    MI_TRY(launch_face_flux(M, c.rho, F->U[0].p, F->U[1].p, F->U[2].p, F->bc.p, F->phi.p, F->s));
    // grad(p) * V of all three components: p does not change during the predictor
    MI_TRY(ensure_ls(M, F->s));
    k_ls_source<<<grid_for(F->n, kBlock), kBlock, 0, F->s>>>(M->dev(), M->ls_minv.p, F->p.p,
                                                             F->src[0].p, F->src[1].p, F->src[2].p);
    MI_HIP(hipGetLastError());
    const bool transient = c.dt > 0.0;
    const bool bdf2 = transient && c.time_scheme == MI_TIME_BDF2 && F->time_step >= 1;
    if (c.momentum_batch)
        return momentum_batched(F, res, transient, bdf2);
    for (int k = 0; k < 3; ++k) {
        MI_TRY(launch_assemble_momentum(M, k, c.mu, c.rho, F->phi.p, F->src[k].p,
                                        transient ? F->dt.p : nullptr, F->Un[k].p,
                                        bdf2 ? F->Unm1[k].p : nullptr, F->U[k].p, c.urfU, F->bc.p,
                                        F->Am->val.p, F->rhs.p, F->s));
        // The matrix values do not depend on the component (only the rhs
        // does): components 1 and 2 rewrite the same values, so the momentum
        // solver keeps the preconditioner it set up for component 0.
        if (k == 0)
            F->Am->values_changed(false);
        if (k == 0) {
            k_take_diag<<<grid_for(F->n, kBlock), kBlock, 0, F->s>>>(F->n, M->pat->diag.p,
                                                                     F->Am->val.p, F->aP.p);
            MI_HIP(hipGetLastError());
        }
        MI_TRY(flow_solve(F, F->sm.get(), F->rhs.p, F->U[k].p));
        MI_TRY(flow_resid(F, F->Am.get(), F->U[k].p, F->rhs.p, F->scal.p + k));
    }
    MI_HIP(hipMemcpyAsync(F->h_scal, F->scal.p, 3 * sizeof(double), hipMemcpyDeviceToHost, F->s));
    MI_HIP(hipStreamSynchronize(F->s));
    for (int k = 0; k < 3; ++k)
        res->mom[k] = F->h_scal[k];
    return MI_OK;
}

// assemble_pressure_correction + solve_pressure_correction_and_correct
// (SimpleAlgorithm.cpp:358-447, 510-519); returns ||rhs||.
int pressure_corrector(mi_flow* F, double* cont)
{
    mi_flow_cfg& c = F->cfg;
    mi_mesh* M = F->M;
    const int pin = (!M->any_dirichlet && F->n > 0) ? 1 : 0;
    const double* den = F->aP.p;
    if (c.algorithm == MI_FLOW_SIMPLEC) {
        // 1 / (aP - sum |a_nb|) in place of 1 / aP (SimpleAlgorithm.cpp:378-387)
        MI_TRY(launch_simplec_denominators(M, F->Am->val.p, F->aP.p, F->den.p, F->s));
        den = F->den.p;
    }
    MI_TRY(launch_assemble_pressure(M, c.rho, den, nullptr, F->phi.p, pin, F->Ap->val.p,
                                    F->rhs.p, F->s));
    F->Ap->values_changed(false);
    MI_TRY(flow_solve(F, F->sp.get(), F->rhs.p, F->pp.p));
    MI_TRY(flow_norm(F, F->rhs.p, F->scal.p + 3));
    MI_TRY(launch_correct_fields(M, c.urfP, c.rho, F->aP.p, F->pp.p, F->p.p, F->U[0].p, F->U[1].p,
                                 F->U[2].p, F->phi.p, F->s));
    MI_HIP(hipMemcpyAsync(F->h_scal + 3, F->scal.p + 3, sizeof(double), hipMemcpyDeviceToHost,
                          F->s));
    MI_HIP(hipStreamSynchronize(F->s));
    *cont = F->h_scal[3];
    return MI_OK;
}

int flow_iterate(mi_flow* F, mi_flow_residuals* res)
{
    std::memset(res, 0, sizeof(*res));
    if (F->n == 0)
        return MI_OK;
    mi_flow_cfg& c = F->cfg;
    MI_TRY(momentum_predictor(F, res));
    if (c.algorithm == MI_FLOW_PISO) {
        // PisoAlgorithm::iterate (PisoAlgorithm.cpp:48-67)
        for (int corr = 0; corr < c.n_correctors; ++corr) {
            MI_TRY(launch_face_flux(F->M, c.rho, F->U[0].p, F->U[1].p, F->U[2].p, F->bc.p,
                                    F->phi.p, F->s));
            for (int no = 0; no <= c.n_nonorth; ++no)
                MI_TRY(pressure_corrector(F, &res->cont));
        }
        return MI_OK;
    }
    // SimpleAlgorithm::iterate (SimpleAlgorithm.cpp:538-540)
    MI_TRY(launch_face_flux(F->M, c.rho, F->U[0].p, F->U[1].p, F->U[2].p, F->bc.p, F->phi.p,
                            F->s));
    return pressure_corrector(F, &res->cont);
}

} // namespace

extern "C" {

MI_API void mi_flow_cfg_default(int algorithm, mi_flow_cfg* c)
{
    if (!c)
        return;
    std::memset(c, 0, sizeof(*c));
    c->algorithm = algorithm;
    c->rho = 1.0;
    c->mu = 1.0e-3;
    if (algorithm == MI_FLOW_SIMPLEC) {
        // SimplecOptions (SimplecAlgorithm.hpp:27-35)
        c->urfU = 0.9;
        c->urfP = 1.0;
        c->dt = 0.0;
        c->time_scheme = MI_TIME_EULER;
        c->n_correctors = 1;
        c->n_nonorth = 0;
    } else if (algorithm == MI_FLOW_PISO) {
        // PisoOptions (PisoAlgorithm.hpp:33-41) and to_simple_opts (:63-74)
        c->urfU = 1.0;
        c->urfP = 1.0;
        c->dt = 1.0e-3;
        c->time_scheme = MI_TIME_BDF2;
        c->n_correctors = 2;
        c->n_nonorth = 0;
    } else {
        // SimpleOptions (SimpleAlgorithm.hpp:43-57)
        c->urfU = 0.7;
        c->urfP = 0.3;
        c->dt = 0.0;
        c->time_scheme = MI_TIME_EULER;
        c->n_correctors = 1;
        c->n_nonorth = 0;
    }
    // LinearSolverConfig's defaults (Solver.hpp:83-90) for the momentum
    // solves; the pressure system is SPD: CG, here with SA-AMG
    mi_solver_cfg_default(&c->momentum);
    c->momentum.kind = MI_SOLVER_GMRES;
    c->momentum.preconditioner = MI_PC_ILU;
    mi_solver_cfg_default(&c->pressure);
    c->pressure.kind = MI_SOLVER_CG;
    c->pressure.preconditioner = MI_PC_AMG;
}

MI_API int mi_flow_create(mi_mesh* M, const double* bc_value_host, const mi_flow_cfg* cfg,
                          mi_flow** out)
{
    clear_error();
    if (!M || !out)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    *out = nullptr;
    MI_TRY(ensure_device());
    auto F = std::make_unique<mi_flow>();
    if (cfg)
        F->cfg = *cfg;
    else
        mi_flow_cfg_default(MI_FLOW_SIMPLE, &F->cfg);
    mi_flow_cfg& c = F->cfg;
    if (c.algorithm != MI_FLOW_SIMPLE && c.algorithm != MI_FLOW_PISO &&
        c.algorithm != MI_FLOW_SIMPLEC)
        return set_error(MI_ERR_BAD_ARG, "unknown algorithm %d", c.algorithm);
    if (!(c.urfU > 0.0) || !(c.dt >= 0.0) || c.n_nonorth < 0)
        return set_error(MI_ERR_BAD_ARG, "bad flow options");
    if (c.time_scheme != MI_TIME_EULER && c.time_scheme != MI_TIME_BDF2)
        return set_error(MI_ERR_BAD_ARG, "unknown time scheme %d", c.time_scheme);
    if (c.algorithm == MI_FLOW_PISO && c.n_correctors < 1)
        c.n_correctors = 1; // PisoAlgorithm.cpp:22-23
    F->M = M;
    F->n = int(M->n_cells);
    F->nf = int(M->n_faces);
    if (hipStreamCreateWithFlags(&F->s, hipStreamNonBlocking) != hipSuccess)
        return set_error(MI_ERR_DEVICE, "hipStreamCreate failed");
    const size_t n = size_t(std::max(F->n, 1)), nf = size_t(std::max(F->nf, 1));
    for (int k = 0; k < 3; ++k) {
        MI_TRY(F->U[k].alloc(n));
        MI_TRY(F->Un[k].alloc(n));
        MI_TRY(F->Unm1[k].alloc(n));
        MI_TRY(F->src[k].alloc(n));
    }
    MI_TRY(F->p.alloc(n));
    MI_TRY(F->aP.alloc(n));
    MI_TRY(F->den.alloc(n));
    if (c.momentum_batch) {
        MI_TRY(F->rhs3.alloc(3 * n));
        MI_TRY(F->x3.alloc(3 * n));
    }
    MI_TRY(F->pp.alloc(n));
    MI_TRY(F->rhs.alloc(n));
    MI_TRY(F->dt.alloc(n));
    MI_TRY(F->phi.alloc(nf));
    MI_TRY(F->partials.alloc(size_t(std::max(kMaxReduceBlocks, kMaxStreamBlocks))));
    MI_TRY(F->scal.alloc(8));
    MI_HIP(hipHostMalloc(reinterpret_cast<void**>(&F->h_scal), 8 * sizeof(double)));
    // zero fields and BDF2 history (build_sparsity: SimpleAlgorithm.cpp:84-96)
    DevBuf<double>* zero[] = {&F->U[0],  &F->U[1],  &F->U[2],    &F->Un[0],   &F->Un[1],
                              &F->Un[2], &F->Unm1[0], &F->Unm1[1], &F->Unm1[2], &F->p,
                              &F->aP,    &F->pp,    &F->phi};
    for (DevBuf<double>* z : zero) {
        k_fill<<<grid_for(int64_t(z->n), kBlock), kBlock, 0, F->s>>>(int(z->n), 0.0, z->p);
        MI_HIP(hipGetLastError());
    }
    k_fill<<<grid_for(F->n, kBlock), kBlock, 0, F->s>>>(F->n, c.dt > 0.0 ? c.dt : 1.0, F->dt.p);
    MI_HIP(hipGetLastError());
    MI_TRY(mesh_bc_values(M, bc_value_host, F->bc, F->s));
    MI_TRY(ensure_ls(M, F->s));
    MI_HIP(hipStreamSynchronize(F->s));
    mi_csr* a = nullptr;
    MI_TRY(mi_mesh_create_matrix(M, &a));
    F->Am.reset(a);
    MI_TRY(mi_mesh_create_matrix(M, &a));
    F->Ap.reset(a);
    mi_solver* sv = nullptr;
    c.momentum.warm_start = 0;
    c.pressure.warm_start = 0;
    MI_TRY(mi_solver_create(&c.momentum, &sv));
    F->sm.reset(sv);
    MI_TRY(mi_solver_create(&c.pressure, &sv));
    F->sp.reset(sv);
    // bind now; the preconditioners are built at the first solve, on real values
    F->sm->A = F->Am.get();
    F->sp->A = F->Ap.get();
    *out = F.release();
    return MI_OK;
}

MI_API int mi_flow_fields(mi_flow* F, double** Ux, double** Uy, double** Uz, double** p,
                          double** phi)
{
    if (!F)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    if (Ux)
        *Ux = F->U[0].p;
    if (Uy)
        *Uy = F->U[1].p;
    if (Uz)
        *Uz = F->U[2].p;
    if (p)
        *p = F->p.p;
    if (phi)
        *phi = F->phi.p;
    return MI_OK;
}

MI_API int mi_flow_iterate(mi_flow* F, mi_flow_residuals* res)
{
    clear_error();
    if (!F)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    MI_TRY(ensure_device());
    mi_flow_residuals local{};
    return flow_iterate(F, res ? res : &local);
}

MI_API int mi_flow_advance_time_step(mi_flow* F, int n_inner, mi_flow_residuals* res)
{
    clear_error();
    if (!F)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    MI_TRY(ensure_device());
    // SimpleAlgorithm::advance_time_step (SimpleAlgorithm.cpp:557-575)
    const size_t bytes = size_t(F->n) * sizeof(double);
    for (int k = 0; k < 3 && F->n > 0; ++k) {
        if (F->time_step >= 1)
            MI_HIP(hipMemcpyAsync(F->Unm1[k].p, F->Un[k].p, bytes, hipMemcpyDeviceToDevice, F->s));
        MI_HIP(hipMemcpyAsync(F->Un[k].p, F->U[k].p, bytes, hipMemcpyDeviceToDevice, F->s));
    }
    mi_flow_residuals local{};
    mi_flow_residuals* r = res ? res : &local;
    for (int k = 0; k < std::max(1, n_inner); ++k)
        MI_TRY(flow_iterate(F, r));
    ++F->time_step;
    return MI_OK;
}

MI_API int mi_flow_destroy(mi_flow* F)
{
    delete F;
    return MI_OK;
}

MI_API int mi_flow_checkpoint_write(mi_flow* F, const char* path, const mi_mesh_soa* mesh,
                                    const mi_mesh_nodes* nodes, const mi_checkpoint_meta* meta,
                                    int flags)
{
    clear_error();
    if (!F || !path || !mesh || !nodes || (flags & ~MI_CHECKPOINT_AMG))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    const AmgHierarchyHost* amg = nullptr;
    if (flags & MI_CHECKPOINT_AMG) {
        amg = solver_amg(F->sp.get());
        if (!amg)
            return set_error(MI_ERR_UNSUPPORTED, "the pressure solver has built no AMG hierarchy");
    }
    if (mesh->n_cells != F->n || mesh->n_faces != F->nf)
        return set_error(MI_ERR_BAD_ARG, "mesh arrays do not match the flow's mesh");
    MI_TRY(ensure_device());
    const size_t n = size_t(F->n);
    std::vector<double> h[4];
    for (int k = 0; k < 4; ++k) {
        h[k].resize(std::max<size_t>(n, 1));
        if (n)
            MI_HIP(hipMemcpyAsync(h[k].data(), k < 3 ? F->U[k].p : F->p.p, n * sizeof(double),
                                  hipMemcpyDeviceToHost, F->s));
    }
    MI_HIP(hipStreamSynchronize(F->s));
    const mi_named_field sc{"p", h[3].data()};
    const mi_named_vector vc{"U", h[0].data(), h[1].data(), h[2].data()};
    MI_TRY(mi_checkpoint_write(path, mesh, nodes, 1, &sc, 1, &vc, meta));
    return amg ? checkpoint_append_amg(path, *amg) : MI_OK;
}

MI_API int mi_flow_checkpoint_read(mi_flow* F, const char* path, mi_checkpoint_meta* meta)
{
    clear_error();
    if (!F || !path)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    mi_checkpoint* c = nullptr;
    MI_TRY(mi_checkpoint_open(path, &c));
    std::unique_ptr<mi_checkpoint, int (*)(mi_checkpoint*)> guard(c, mi_checkpoint_close);
    int64_t nc = 0;
    MI_TRY(mi_checkpoint_info(c, meta, &nc, nullptr, nullptr));
    if (nc != F->n)
        return set_error(MI_ERR_BAD_ARG, "checkpoint has %lld cells, the flow %d",
                         (long long)nc, F->n);
    const size_t n = size_t(F->n);
    std::vector<double> h[4];
    for (auto& v : h)
        v.resize(std::max<size_t>(n, 1));
    int64_t cnt = 0;
    MI_TRY(mi_checkpoint_scalar(c, "p", h[3].data(), int64_t(n), &cnt));
    if (cnt != int64_t(n))
        return set_error(MI_ERR_BAD_ARG, "field p has %lld values", (long long)cnt);
    MI_TRY(mi_checkpoint_vector(c, "U", h[0].data(), h[1].data(), h[2].data(), int64_t(n), &cnt));
    if (cnt != int64_t(n))
        return set_error(MI_ERR_BAD_ARG, "field U has %lld values", (long long)cnt);
    for (int k = 0; k < 4 && n; ++k)
        MI_HIP(hipMemcpyAsync(k < 3 ? F->U[k].p : F->p.p, h[k].data(), n * sizeof(double),
                              hipMemcpyHostToDevice, F->s));
    MI_HIP(hipStreamSynchronize(F->s));
    // no time history survives a restart: the next step starts it again
    F->time_step = 0;
    // a cached pressure hierarchy replaces the first setup of a solver that
    // keeps its preconditioner
    const mi_solver_cfg& pc = F->cfg.pressure;
    if (const AmgHierarchyHost* amg = checkpoint_amg(c))
        if (pc.preconditioner == MI_PC_AMG && pc.keep_preconditioner)
            MI_TRY(solver_adopt_amg(F->sp.get(), *amg));
    return MI_OK;
}

MI_API int mi_ls_gradient(mi_mesh* M, const double* phi, double* gx, double* gy, double* gz)
{
    clear_error();
    if (!M || (M->n_cells > 0 && (!phi || !gx || !gy || !gz)))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    MI_TRY(launch_ls_gradient(M, phi, gx, gy, gz, s));
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

MI_API int mi_correct_fields(mi_mesh* M, double urfP, double rho, const double* aP,
                             const double* pprime, double* p, double* Ux, double* Uy, double* Uz,
                             double* phi)
{
    clear_error();
    if (!M || (M->n_cells > 0 && (!aP || !pprime || !p || !Ux || !Uy || !Uz)))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    MI_TRY(launch_correct_fields(M, urfP, rho, aP, pprime, p, Ux, Uy, Uz, phi, s));
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

} // extern "C"
