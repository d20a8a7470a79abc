// assembly.hip — finite-volume assembly of the pressure-correction and
// momentum systems from the SoA mesh, and the face mass flux.
//
// Reference: SimpleAlgorithm::build_sparsity / compute_face_fluxes /
// assemble_momentum / assemble_pressure_correction
// (src/solver/SimpleAlgorithm.cpp:49-98, 201-237, 240-355, 358-447).
//
// The reference loops over faces and scatters into four CSR entries per face
// with a linear column search.  Here one thread owns one cell (one CSR row):
// it walks the cell's faces in ascending face id — the order in which the
// reference's face loop touches that row — and gathers every contribution, so
// there are no atomics and each row's sums are formed in the reference's
// order.  With floating-point contraction off the coefficients are the same
// IEEE operations as the reference's, so the assembled values are bit-exact.
#include "mesh.hpp"
#include "host_threads.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>

namespace mi
{

#pragma clang fp contract(off)

// Pressure-correction row of cell c (src/solver/SimpleAlgorithm.cpp:358-447,
// SIMPLE denominators).
__global__ __launch_bounds__(kBlock) void k_assemble_pressure(MeshDev m, double rho,
                                                              const double* __restrict__ aP,
                                                              const double* __restrict__ gamma,
                                                              const double* __restrict__ phi,
                                                              int pin, double* __restrict__ val,
                                                              double* __restrict__ rhs)
{
    const int c = int(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= m.n_cells)
        return;
    // direct: the slots write distinct off-diagonal entries, each stored as
    // 0 - coef (the zeroed entry minus coef) without reading it back
    const int cs = m.ss_cnt[c];
    const bool direct = cs >= 0;
    const int cnt = direct ? cs : -cs;
    if (!direct)
        for (int k = m.rp[c]; k < m.rp[c + 1]; ++k)
            val[k] = 0.0;
    double diag = 0.0, b = 0.0;
    const int q0 = m.ss_off[c / kSlotSlice] + c % kSlotSlice;
    for (int j = 0; j < cnt; ++j) {
        const int q = q0 + j * kSlotSlice;
        const int f = m.ss_face[q];
        const int oth = m.ss_oth[q];
        const int tag = m.ss_tag[q];
        const bool own = (tag & kSlotOwner) != 0;
        const double Amag2 = m.ss_s2[q];
        const double ddotA = m.ss_ds[q];
        const double g = gamma ? gamma[f] : 1.0;
        if (tag & kSlotPeriodic) {
            // the owner's half of the coupling with the twin's cell (:411-419)
            const double dO = aP ? fmax(aP[c], 1e-30) : 1.0;
            const double coef = rho * g * Amag2 / fmax(dO * ddotA, 1e-30);
            diag += coef;
            if (oth == c) {
                diag += -coef; // a cell that is its own twin: the same entry
            } else {
                const int e = m.ss_slot[q];
                val[e] = (direct ? 0.0 : val[e]) - coef;
            }
        } else if (oth >= 0) {
            double aPf = 1.0;
            if (aP) {
                const double dO = fmax(aP[own ? c : oth], 1e-30);
                const double dN = fmax(aP[own ? oth : c], 1e-30);
                aPf = 0.5 * (dO + dN);
            }
            const double coef = rho * g * Amag2 / fmax(aPf * ddotA, 1e-30);
            diag += coef;
            const int e = m.ss_slot[q];
            val[e] = (direct ? 0.0 : val[e]) - coef;
        } else if ((tag & 15) - 1 == BC_PRESSURE) {
            const double dO = aP ? fmax(aP[c], 1e-30) : 1.0;
            const double coef = rho * g * Amag2 / fmax(dO * ddotA, 1e-30);
            diag += coef;
        }
        if (phi) {
            if (own)
                b -= phi[f];
            else
                b += phi[f];
        }
    }
    const int kd = m.diag[c];
    if (kd >= 0) {
        double v = (direct ? 0.0 : val[kd]) + diag;
        if (pin && c == 0)
            v *= 1.0e6;
        val[kd] = v;
    }
    if (pin && c == 0)
        b = 0.0;
    rhs[c] = b;
}

// The sliced copies of the slot arrays (MeshDev::ss_*), once per mesh:
// element q takes slot smap[q] of the cell-order arrays (-1: padding).
__global__ __launch_bounds__(kBlock) void k_slice_slots(MeshDev m, const int* __restrict__ smap,
                                                        int64_t total, int* __restrict__ face,
                                                        int* __restrict__ slot, int* __restrict__ oth,
                                                        int8_t* __restrict__ tag,
                                                        double* __restrict__ s2,
                                                        double* __restrict__ ds)
{
    for (int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x; q < total;
         q += int64_t(gridDim.x) * kBlock) {
        const int k = smap[q];
        face[q] = k >= 0 ? m.cf_face[k] : -1;
        slot[q] = k >= 0 ? m.cf_slot[k] : -1;
        oth[q] = k >= 0 ? m.sl_oth[k] : -1;
        tag[q] = k >= 0 ? m.sl_tag[k] : int8_t(0);
        s2[q] = k >= 0 ? m.sl_s2[k] : 0.0;
        ds[q] = k >= 0 ? m.sl_ds[k] : 0.0;
    }
}

// Geometry of every cell-face slot (MeshDev::sl_s2 / sl_ds), once per mesh:
// |S|^2 and |d . S| with the operations the assembly kernels would apply to
// the face and cell arrays, so the kernels' arithmetic is unchanged.
__global__ __launch_bounds__(kBlock) void k_slot_geom(MeshDev m, double* __restrict__ s2,
                                                      double* __restrict__ ds)
{
    const int c = int(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= m.n_cells)
        return;
    for (int k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
        const int f = m.cf_face[k];
        const int o = m.owner[f];
        const int nb = m.nbr[f];
        const double Ax = m.ax[f], Ay = m.ay[f], Az = m.az[f];
        double dx, dy, dz;
        if (nb >= 0) {
            dx = m.ccx[nb] - m.ccx[o];
            dy = m.ccy[nb] - m.ccy[o];
            dz = m.ccz[nb] - m.ccz[o];
        } else {
            dx = m.fcx[f] - m.ccx[o];
            dy = m.fcy[f] - m.ccy[o];
            dz = m.fcz[f] - m.ccz[o];
        }
        s2[k] = Ax * Ax + Ay * Ay + Az * Az;
        ds[k] = fabs(dx * Ax + dy * Ay + dz * Az);
    }
}

// dJ/dgamma_f = -lambda^T (dA/dgamma_f) x for the pressure matrix of
// k_assemble_pressure: an interior face adds k_f (e_o - e_n)(e_o - e_n)^T, a
// pressure-boundary face k_f e_o e_o^T, a periodic face k_f e_o (e_o - e_t)^T
// (t the twin's cell), with k_f = gamma_f * G_f; the pinned
// cell's diagonal carries the factor 1e6.  The rhs does not depend on gamma.
// (Gradient = dJ/dalpha - (dR/dalpha)^T psi: src/adjoint/DiscreteAdjoint.cpp:170-182,
// with R = A(gamma) x - b.)
//
// The face loop streams per-face data only: owner and neighbour in a cell
// numbering that follows the face list (fg_o, fg_n), and |S|^2, |d . S|
// precomputed once per mesh (k_fg_geom) with the operations the coefficient
// of k_assemble_pressure uses; x, lambda and aP are gathered into that
// numbering once per call (k_fg_gather), so the cell reads of neighbouring
// faces share cache lines even on a randomly numbered mesh.
__global__ __launch_bounds__(kBlock) void k_fg_geom(MeshDev m, const int* __restrict__ inv,
                                                    int* __restrict__ fo, int* __restrict__ fn,
                                                    double* __restrict__ s2,
                                                    double* __restrict__ ds)
{
    const int f = int(blockIdx.x) * kBlock + threadIdx.x;
    if (f >= m.n_faces)
        return;
    const int o = m.owner[f];
    const int nb = m.nbr[f];
    const double Ax = m.ax[f], Ay = m.ay[f], Az = m.az[f];
    const double Amag2 = Ax * Ax + Ay * Ay + Az * Az;
    fo[f] = inv[o];
    double d = 0.0;
    int tag = -2;
    if (nb >= 0) {
        const double dx = m.ccx[nb] - m.ccx[o];
        const double dy = m.ccy[nb] - m.ccy[o];
        const double dz = m.ccz[nb] - m.ccz[o];
        d = fabs(dx * Ax + dy * Ay + dz * Az);
        tag = inv[nb];
    } else if (m.ftwin[f] >= 0 || m.fbc[f] == BC_PRESSURE) {
        const double dx = m.fcx[f] - m.ccx[o];
        const double dy = m.fcy[f] - m.ccy[o];
        const double dz = m.fcz[f] - m.ccz[o];
        d = fabs(dx * Ax + dy * Ay + dz * Az);
        tag = m.ftwin[f] >= 0 ? -3 - inv[m.ftwin[f]] : -1;
    }
    fn[f] = tag;
    s2[f] = Amag2;
    ds[f] = d;
}

// xr[i] = x[perm[i]], lr[i] = lam[perm[i]], and aP when given
__global__ __launch_bounds__(kBlock) void k_fg_gather(int n, const int* __restrict__ perm,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ lam,
                                                      const double* __restrict__ aP,
                                                      double* __restrict__ xr,
                                                      double* __restrict__ lr,
                                                      double* __restrict__ ar)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const int c = perm[i];
    xr[i] = x[c];
    lr[i] = lam[c];
    if (aP)
        ar[i] = aP[c];
}

__global__ __launch_bounds__(kBlock) void k_face_gradient(int n_faces, const int* __restrict__ fo,
                                                          const int* __restrict__ fn,
                                                          const double* __restrict__ s2,
                                                          const double* __restrict__ ds,
                                                          double rho, const double* __restrict__ aP,
                                                          int pin_cell,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ lam,
                                                          double* __restrict__ grad)
{
    const int f = int(blockIdx.x) * kBlock + threadIdx.x;
    if (f >= n_faces)
        return;
    const int o = fo[f];
    const int nb = fn[f];
    const double Amag2 = s2[f];
    const double ddotA = ds[f];
    const double so = (o == pin_cell) ? 1.0e6 : 1.0;
    double g = 0.0;
    if (nb >= 0) {
        double aPf = 1.0;
        if (aP)
            aPf = 0.5 * (fmax(aP[o], 1e-30) + fmax(aP[nb], 1e-30));
        const double G = rho * Amag2 / fmax(aPf * ddotA, 1e-30);
        const double sn = (nb == pin_cell) ? 1.0e6 : 1.0;
        const double lo = lam[o], ln = lam[nb], xo = x[o], xn = x[nb];
        const double t = (lo - ln) * (xo - xn) + (so - 1.0) * lo * xo + (sn - 1.0) * ln * xn;
        g = -G * t;
    } else if (nb == -1) {
        const double dO = aP ? fmax(aP[o], 1e-30) : 1.0;
        const double G = rho * Amag2 / fmax(dO * ddotA, 1e-30);
        g = -G * so * lam[o] * x[o];
    } else if (nb <= -3) {
        // periodic: G (e_o e_o^T - e_o e_t^T), t the twin's cell
        const int t = -3 - nb;
        const double dO = aP ? fmax(aP[o], 1e-30) : 1.0;
        const double G = rho * Amag2 / fmax(dO * ddotA, 1e-30);
        g = -G * lam[o] * (so * x[o] - x[t]);
    }
    grad[f] = g;
}

// Momentum row of cell c, one velocity component, first-order upwind +
// diffusion (src/solver/SimpleAlgorithm.cpp:240-355).  pgradV (nullable) is
// the cell's grad(p)_comp * V_c, subtracted from the rhs (:307-315).  With dt
// the time term (:317-336): implicit Euler on Uold, or BDF2 on Uold and Unm1
// when Unm1 is given.
__global__ __launch_bounds__(kBlock) void k_assemble_momentum(
    MeshDev m, int comp, double mu, double rho, const double* __restrict__ phi,
    const double* __restrict__ pgradV, const double* __restrict__ dt,
    const double* __restrict__ Uold, const double* __restrict__ Unm1,
    const double* __restrict__ Ucomp, double urf,
    const double* __restrict__ bcValue, double* __restrict__ val, double* __restrict__ rhs)
{
    const int c = int(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= m.n_cells)
        return;
    // direct: as in k_assemble_pressure
    const int cs = m.ss_cnt[c];
    const bool direct = cs >= 0;
    const int cnt = direct ? cs : -cs;
    if (!direct)
        for (int k = m.rp[c]; k < m.rp[c + 1]; ++k)
            val[k] = 0.0;
    double diag = 0.0, b = 0.0;
    const int q0 = m.ss_off[c / kSlotSlice] + c % kSlotSlice;
    for (int j = 0; j < cnt; ++j) {
        const int q = q0 + j * kSlotSlice;
        const int f = m.ss_face[q];
        const int oth = m.ss_oth[q];
        const int tag = m.ss_tag[q];
        const double Amag2 = m.ss_s2[q];
        const double ddotA = m.ss_ds[q];
        const double Ff = phi[f];
        const double ConvO = fmax(Ff, 0.0);
        const double ConvN = fmax(-Ff, 0.0);
        if (oth >= 0) {
            // interior faces and periodic faces (the owner's half of the
            // coupling with the twin's cell, :279-285)
            const double D = mu * Amag2 / fmax(ddotA, 1e-30);
            if ((tag & kSlotPeriodic) && oth == c) {
                diag += ConvO + D;
                diag += -(ConvN + D);
            } else if (tag & kSlotOwner) {
                diag += ConvO + D;
                const int e = m.ss_slot[q];
                val[e] = (direct ? 0.0 : val[e]) + -(ConvN + D);
            } else {
                diag += ConvN + D;
                const int e = m.ss_slot[q];
                val[e] = (direct ? 0.0 : val[e]) + -(ConvO + D);
            }
        } else {
            const int bc = (tag & 15) - 1;
            const double D = mu * Amag2 / fmax(ddotA, 1e-30);
            if (bc == BC_WALL || bc == BC_VELOCITY_INLET) {
                const double Ub = (bc == BC_WALL) ? 0.0 : bcValue[m.fzone[f] * 3 + comp];
                diag += ConvO + D;
                b += (ConvN + D) * Ub;
            } else if (bc == BC_SYMMETRY) {
            } else {
                diag += ConvO;
            }
        }
    }
    if (pgradV)
        b -= pgradV[c];
    const int kd = m.diag[c];
    double a = (kd >= 0) ? (direct ? 0.0 : val[kd]) + diag : diag;
    if (dt && kd >= 0) {
        const double Vrho_dt = rho * m.vol[c] / dt[c];
        const double un = Uold ? Uold[c] : 0.0;
        if (Unm1) {
            a += 1.5 * Vrho_dt;
            b += (2.0 * un - 0.5 * Unm1[c]) * Vrho_dt;
        } else {
            a += Vrho_dt;
            b += un * Vrho_dt;
        }
    }
    if (kd >= 0) {
        const double scaled = a / urf;
        b += (scaled - a) * (Ucomp ? Ucomp[c] : 0.0);
        val[kd] = scaled;
    }
    rhs[c] = b;
}

// phi_f = rho * (U_f . S_f) (src/solver/SimpleAlgorithm.cpp:201-237).
__global__ __launch_bounds__(kBlock) void k_face_flux(MeshDev m, double rho,
                                                      const double* __restrict__ Ux,
                                                      const double* __restrict__ Uy,
                                                      const double* __restrict__ Uz,
                                                      const double* __restrict__ bcValue,
                                                      double* __restrict__ phi)
{
    const int f = int(blockIdx.x) * kBlock + threadIdx.x;
    if (f >= m.n_faces)
        return;
    const int o = m.owner[f], nb = m.nbr[f];
    double ux, uy, uz;
    const int tw = nb < 0 ? m.ftwin[f] : -1;
    if (nb >= 0) {
        ux = 0.5 * (Ux[o] + Ux[nb]);
        uy = 0.5 * (Uy[o] + Uy[nb]);
        uz = 0.5 * (Uz[o] + Uz[nb]);
    } else if (tw >= 0) {
        ux = 0.5 * (Ux[o] + Ux[tw]); // periodic (:214-218)
        uy = 0.5 * (Uy[o] + Uy[tw]);
        uz = 0.5 * (Uz[o] + Uz[tw]);
    } else {
        const int bc = m.fbc[f];
        if (bc == BC_VELOCITY_INLET) {
            const int z = m.fzone[f];
            ux = bcValue[z * 3 + 0];
            uy = bcValue[z * 3 + 1];
            uz = bcValue[z * 3 + 2];
        } else if (bc == BC_WALL || bc == BC_SYMMETRY) {
            ux = uy = uz = 0.0;
        } else {
            ux = Ux[o];
            uy = Uy[o];
            uz = Uz[o];
        }
    }
    phi[f] = rho * (ux * m.ax[f] + uy * m.ay[f] + uz * m.az[f]);
}

// SIMPLEC denominator of cell c (src/solver/SimpleAlgorithm.cpp:378-387,
// 400-409): max(aP_c - sum of |a| over row c's off-diagonal entries of the
// momentum matrix, in row order, 1e-30).
__global__ __launch_bounds__(kBlock) void k_simplec_den(int n, const int* __restrict__ rp,
                                                       const int* __restrict__ ci,
                                                       const double* __restrict__ amom,
                                                       const double* __restrict__ aP,
                                                       double* __restrict__ den)
{
    const int c = int(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n)
        return;
    double sumOff = 0.0;
    for (int k = rp[c]; k < rp[c + 1]; ++k)
        if (ci[k] != c)
            sumOff += fabs(amom[k]);
    den[c] = fmax(aP[c] - sumOff, 1e-30);
}

#pragma clang fp contract(on)

int launch_simplec_denominators(const mi_mesh* M, const double* amom, const double* aP,
                                double* den, hipStream_t s)
{
    if (M->n_cells > 0) {
        k_simplec_den<<<grid_for(M->n_cells, kBlock), kBlock, 0, s>>>(
            int(M->n_cells), M->pat->rp.p, M->pat->ci.p, amom, aP, den);
        MI_HIP(hipGetLastError());
    }
    return MI_OK;
}

int launch_assemble_pressure(const mi_mesh* M, double rho, const double* aP, const double* gamma,
                             const double* phi, int pin, double* val, double* rhs, hipStream_t s)
{
    if (M->n_cells > 0) {
        k_assemble_pressure<<<grid_for(M->n_cells, kBlock), kBlock, 0, s>>>(
            M->dev(), rho, aP, gamma, phi, pin, val, rhs);
        MI_HIP(hipGetLastError());
    }
    return MI_OK;
}

int launch_assemble_momentum(const mi_mesh* M, int comp, double mu, double rho, const double* phi,
                             const double* pgradV, const double* dt, const double* Uold,
                             const double* Unm1, const double* Ucomp, double urf,
                             const double* bc_dev, double* val, double* rhs, hipStream_t s)
{
    if (M->n_cells > 0) {
        k_assemble_momentum<<<grid_for(M->n_cells, kBlock), kBlock, 0, s>>>(
            M->dev(), comp, mu, rho, phi, pgradV, dt, Uold, Unm1, Ucomp, urf, bc_dev, val, rhs);
        MI_HIP(hipGetLastError());
    }
    return MI_OK;
}

int launch_face_flux(const mi_mesh* M, double rho, const double* Ux, const double* Uy,
                     const double* Uz, const double* bc_dev, double* phi, hipStream_t s)
{
    if (M->n_faces > 0) {
        k_face_flux<<<grid_for(M->n_faces, kBlock), kBlock, 0, s>>>(M->dev(), rho, Ux, Uy, Uz,
                                                                    bc_dev, phi);
        MI_HIP(hipGetLastError());
    }
    return MI_OK;
}

int mesh_bc_values(const mi_mesh* M, const double* h, DevBuf<double>& d, hipStream_t s)
{
    std::vector<double> v(size_t(std::max(M->n_zones, 1)) * 3, 0.0);
    if (h)
        std::memcpy(v.data(), h, size_t(M->n_zones) * 3 * sizeof(double));
    return d.upload(v.data(), v.size(), s);
}

} // namespace mi

using namespace mi;

namespace
{

// Face-gradient state of a mesh (see mi_mesh::fg_*), built once.
int face_gradient_setup(mi_mesh* M, hipStream_t s)
{
    if (M->fg_o.p)
        return MI_OK;
    const int n = int(M->n_cells), F = int(M->n_faces);
    std::vector<int> own(static_cast<size_t>(F)), nbr(static_cast<size_t>(F));
    MI_HIP(hipMemcpyAsync(own.data(), M->owner.p, sizeof(int) * size_t(F), hipMemcpyDeviceToHost, s));
    MI_HIP(hipMemcpyAsync(nbr.data(), M->nbr.p, sizeof(int) * size_t(F), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    // cells in the order the face list first meets them, then any it never meets
    std::vector<int> inv(size_t(n), -1), perm;
    perm.reserve(size_t(n));
    auto see = [&](int c) {
        if (c >= 0 && inv[size_t(c)] < 0) {
            inv[size_t(c)] = int(perm.size());
            perm.push_back(c);
        }
    };
    for (int f = 0; f < F; ++f) {
        see(own[size_t(f)]);
        see(nbr[size_t(f)]);
    }
    for (int c = 0; c < n; ++c)
        see(c);
    DevBuf<int> d_inv;
    MI_TRY(d_inv.upload(inv.data(), inv.size(), s));
    MI_TRY(M->fg_perm.upload(perm.data(), perm.size(), s));
    MI_TRY(M->fg_o.alloc(size_t(F)));
    MI_TRY(M->fg_n.alloc(size_t(F)));
    MI_TRY(M->fg_s2.alloc(size_t(F)));
    MI_TRY(M->fg_ds.alloc(size_t(F)));
    k_fg_geom<<<grid_for(F, kBlock), kBlock, 0, s>>>(M->dev(), d_inv.p, M->fg_o.p, M->fg_n.p,
                                                     M->fg_s2.p, M->fg_ds.p);
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s));
    M->fg_pin = n > 0 ? inv[0] : -1;
    return MI_OK;
}

int bc_of(const mi_mesh* M, uint32_t zone)
{
    // A zone with no entry behaves as the reference's find_bc == nullptr: the
    // fall-through (outflow) branch of every face loop.
    if (int64_t(zone) >= M->n_zones)
        return BC_OUTFLOW;
    return M->bc_of_zone[zone];
}

template <class T> int up(mi::DevBuf<T>& d, const T* h, int64_t n, hipStream_t s)
{
    if (n > 0 && !h)
        return set_error(MI_ERR_BAD_ARG, "null mesh array");
    return d.upload(h, size_t(n), s);
}

// SimpleAlgorithm::build_periodic_pairs (src/solver/SimpleAlgorithm.cpp:101-186)
// on the host: the periodic zones in ascending zone id are paired two by two;
// t = mean centroid of the second zone's boundary faces - mean of the
// first's; every face fa of the first zone (face order) takes the first face
// fb of the second zone with the least |c_fb - (c_fa + t)|^2, and
// twin[fa] = fb, twin[fb] = fa (a later fa overwrites twin[fb]).  The
// nearest-face searches run on host threads; the pairs are written in face
// order afterwards.
void periodic_pairs(const mi_mesh_soa* m, const std::vector<int>& bc_of_zone,
                    std::vector<int>& twin)
{
    const int64_t nF = m->n_faces;
    twin.assign(size_t(nF), -1);
    std::vector<int> zones;
    for (size_t z = 0; z < bc_of_zone.size(); ++z)
        if (bc_of_zone[z] == BC_PERIODIC)
            zones.push_back(int(z));
    if (zones.size() < 2 || !m->boundaryZone || !m->faceCx)
        return;
    for (size_t zi = 0; zi + 1 < zones.size(); zi += 2) {
        std::vector<int> A, B;
        for (int64_t f = 0; f < nF; ++f) {
            if (m->neighbor[f] != UINT64_MAX)
                continue;
            if (int64_t(m->boundaryZone[f]) == zones[zi])
                A.push_back(int(f));
            else if (int64_t(m->boundaryZone[f]) == zones[zi + 1])
                B.push_back(int(f));
        }
        if (A.empty() || B.empty())
            continue;
        double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
        for (int f : A) {
            ax += m->faceCx[f];
            ay += m->faceCy[f];
            az += m->faceCz[f];
        }
        for (int f : B) {
            bx += m->faceCx[f];
            by += m->faceCy[f];
            bz += m->faceCz[f];
        }
        const double ia = 1.0 / double(A.size()), ib = 1.0 / double(B.size());
        const double tx = bx * ib - ax * ia, ty = by * ib - ay * ia, tz = bz * ib - az * ia;
        std::vector<int> best(A.size(), -1);
        host_parallel_for(int(A.size()), host_threads(), [&](int b0, int e0, int) {
            for (int i = b0; i < e0; ++i) {
                const int fa = A[size_t(i)];
                const double px = m->faceCx[fa] + tx, py = m->faceCy[fa] + ty,
                             pz = m->faceCz[fa] + tz;
                double bestd2 = std::numeric_limits<double>::max();
                for (int fb : B) {
                    const double d2 = (m->faceCx[fb] - px) * (m->faceCx[fb] - px) +
                                      (m->faceCy[fb] - py) * (m->faceCy[fb] - py) +
                                      (m->faceCz[fb] - pz) * (m->faceCz[fb] - pz);
                    if (d2 < bestd2) {
                        bestd2 = d2;
                        best[size_t(i)] = fb;
                    }
                }
            }
        });
        for (size_t i = 0; i < A.size(); ++i)
            if (best[i] >= 0) {
                twin[size_t(A[i])] = best[i];
                twin[size_t(best[i])] = A[i];
            }
    }
}

int mesh_create(const mi_mesh_soa* m, const int32_t* bc_of_zone, int n_zones,
                const int32_t* twin_faces, mi_mesh** out)
{
    clear_error();
    if (!m || !out || n_zones < 0 || (n_zones > 0 && !bc_of_zone))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    *out = nullptr;
    MI_TRY(ensure_device());
    const int64_t nC = m->n_cells, nF = m->n_faces;
    if (nC < 0 || nF < 0 || nC > INT32_MAX - 1 || nF > INT32_MAX - 1)
        return set_error(MI_ERR_BAD_ARG, "mesh too large for int32 ids");
    if (nC > 0 && (!m->cellFaceOffsets || !m->cellFaceIndices || !m->owner || !m->neighbor))
        return set_error(MI_ERR_BAD_ARG, "null connectivity");
    auto M = std::make_unique<mi_mesh>();
    M->n_cells = nC;
    M->n_faces = nF;
    M->n_zones = n_zones;
    M->bc_of_zone.assign(bc_of_zone, bc_of_zone + n_zones);
    hipStream_t s = default_stream();

    // periodic twins: the caller's table (SimpleAlgorithm::set_periodic_pairs)
    // or the pairing of the periodic zones; only boundary faces keep one
    if (twin_faces) {
        M->twin_face.assign(twin_faces, twin_faces + nF);
        for (int64_t f = 0; f < nF; ++f)
            if (M->twin_face[size_t(f)] < -1 || M->twin_face[size_t(f)] >= nF)
                return set_error(MI_ERR_BAD_ARG, "face %lld: twin out of range", (long long)f);
    } else {
        periodic_pairs(m, M->bc_of_zone, M->twin_face);
    }
    // face arrays, ids narrowed to int32 (kBoundaryCell -> -1)
    std::vector<int> own(static_cast<size_t>(nF)), nbr(static_cast<size_t>(nF)),
        zone(static_cast<size_t>(nF)), ftwin(static_cast<size_t>(nF), -1);
    std::vector<int8_t> fbc(static_cast<size_t>(nF));
    for (int64_t f = 0; f < nF; ++f) {
        if (m->owner[f] >= uint64_t(nC))
            return set_error(MI_ERR_BAD_ARG, "face %lld: owner out of range", (long long)f);
        own[size_t(f)] = int(m->owner[f]);
        const uint64_t nb = m->neighbor[f];
        if (nb != UINT64_MAX && nb >= uint64_t(nC))
            return set_error(MI_ERR_BAD_ARG, "face %lld: neighbour out of range", (long long)f);
        nbr[size_t(f)] = (nb == UINT64_MAX) ? -1 : int(nb);
        const uint32_t z = m->boundaryZone ? m->boundaryZone[f] : 0u;
        zone[size_t(f)] = int(std::min<uint32_t>(z, uint32_t(INT32_MAX)));
        if (nb == UINT64_MAX) {
            const int bc = bc_of(M.get(), z);
            fbc[size_t(f)] = int8_t(bc);
            if (bc == BC_PRESSURE && M->twin_face[size_t(f)] < 0)
                M->any_dirichlet = true;
        } else {
            fbc[size_t(f)] = -1;
        }
    }
    for (int64_t f = 0; f < nF; ++f) {
        const int t = M->twin_face[size_t(f)];
        if (t >= 0 && nbr[size_t(f)] < 0)
            ftwin[size_t(f)] = own[size_t(t)];
    }
    // CSR pattern: row c = {c} U face neighbours, sorted, unique
    // (src/solver/SimpleAlgorithm.cpp:63-83); per-cell faces sorted by id and
    // the CSR slot of the cell across each interior face.
    std::vector<int> rp(size_t(nC) + 1, 0), ci;
    std::vector<int> cf_off(size_t(nC) + 1, 0), cf_face, cf_slot, sl_oth;
    std::vector<int8_t> sl_tag;
    ci.reserve(size_t(nC) * 7);
    cf_face.reserve(size_t(nC) * 6);
    cf_slot.reserve(size_t(nC) * 6);
    std::vector<int> nb_list, faces;
    for (int64_t c = 0; c < nC; ++c) {
        const int beg = m->cellFaceOffsets[c], end = m->cellFaceOffsets[c + 1];
        if (end < beg)
            return set_error(MI_ERR_BAD_ARG, "cellFaceOffsets decreases at cell %lld", (long long)c);
        nb_list.assign(1, int(c));
        faces.clear();
        for (int k = beg; k < end; ++k) {
            const uint64_t f = m->cellFaceIndices[k];
            if (f >= uint64_t(nF))
                return set_error(MI_ERR_BAD_ARG, "cell %lld: face index out of range", (long long)c);
            faces.push_back(int(f));
            const int o = own[size_t(f)], nb = nbr[size_t(f)];
            const int oth = (o == int(c)) ? nb : o;
            if (oth >= 0)
                nb_list.push_back(oth);
            else if (ftwin[size_t(f)] >= 0)
                nb_list.push_back(ftwin[size_t(f)]);
        }
        std::sort(nb_list.begin(), nb_list.end());
        nb_list.erase(std::unique(nb_list.begin(), nb_list.end()), nb_list.end());
        const int row0 = int(ci.size());
        ci.insert(ci.end(), nb_list.begin(), nb_list.end());
        rp[size_t(c) + 1] = int(ci.size());
        std::sort(faces.begin(), faces.end());
        for (int f : faces) {
            cf_face.push_back(f);
            const int o = own[size_t(f)], nb = nbr[size_t(f)];
            const bool periodic = nb < 0 && ftwin[size_t(f)] >= 0;
            const int oth = periodic ? ftwin[size_t(f)] : (o == int(c)) ? nb : o;
            sl_oth.push_back(oth);
            sl_tag.push_back(int8_t(((o == int(c)) ? kSlotOwner : 0) +
                                    (periodic ? kSlotPeriodic : 0) + (int(fbc[size_t(f)]) + 1)));
            int slot = -1;
            if (oth >= 0)
                slot = row0 + int(std::lower_bound(nb_list.begin(), nb_list.end(), oth) -
                                  nb_list.begin());
            cf_slot.push_back(slot);
        }
        cf_off[size_t(c) + 1] = int(cf_face.size());
    }
    MI_TRY(make_pattern(nC, int64_t(ci.size()), rp.data(), ci.data(), M->pat, s));
    MI_TRY(up(M->owner, own.data(), nF, s));
    MI_TRY(up(M->nbr, nbr.data(), nF, s));
    MI_TRY(up(M->fzone, zone.data(), nF, s));
    MI_TRY(up(M->fbc, fbc.data(), nF, s));
    MI_TRY(up(M->ftwin, ftwin.data(), nF, s));
    MI_TRY(up(M->ax, m->areaX, nF, s));
    MI_TRY(up(M->ay, m->areaY, nF, s));
    MI_TRY(up(M->az, m->areaZ, nF, s));
    MI_TRY(up(M->fcx, m->faceCx, nF, s));
    MI_TRY(up(M->fcy, m->faceCy, nF, s));
    MI_TRY(up(M->fcz, m->faceCz, nF, s));
    MI_TRY(up(M->ccx, m->cellCx, nC, s));
    MI_TRY(up(M->ccy, m->cellCy, nC, s));
    MI_TRY(up(M->ccz, m->cellCz, nC, s));
    MI_TRY(up(M->vol, m->volume, nC, s));
    MI_TRY(up(M->cf_off, cf_off.data(), nC + 1, s));
    MI_TRY(up(M->cf_face, cf_face.data(), int64_t(cf_face.size()), s));
    MI_TRY(up(M->cf_slot, cf_slot.data(), int64_t(cf_slot.size()), s));
    MI_TRY(up(M->sl_oth, sl_oth.data(), int64_t(sl_oth.size()), s));
    MI_TRY(up(M->sl_tag, sl_tag.data(), int64_t(sl_tag.size()), s));
    MI_TRY(M->sl_s2.alloc(std::max<size_t>(cf_face.size(), 1)));
    MI_TRY(M->sl_ds.alloc(std::max<size_t>(cf_face.size(), 1)));
    if (nC > 0) {
        k_slot_geom<<<grid_for(nC, kBlock), kBlock, 0, s>>>(M->dev(), M->sl_s2.p, M->sl_ds.p);
        MI_HIP(hipGetLastError());
    }
    // sliced slot arrays: slices of kSlotSlice cells padded to the slice's
    // longest cell
    {
        const int64_t nsl = (nC + kSlotSlice - 1) / kSlotSlice;
        std::vector<int> soff(size_t(nsl) + 1, 0);
        std::vector<int> cnt(static_cast<size_t>(nC), 0);
        int64_t tot = 0;
        for (int64_t sl = 0; sl < nsl; ++sl) {
            int w = 0;
            const int64_t c1 = std::min<int64_t>(nC, (sl + 1) * kSlotSlice);
            for (int64_t c = sl * kSlotSlice; c < c1; ++c) {
                cnt[size_t(c)] = cf_off[size_t(c) + 1] - cf_off[size_t(c)];
                w = std::max(w, cnt[size_t(c)]);
            }
            // a cell whose slots write distinct off-diagonal entries of its row
            // (every entry but the diagonal, by construction of the pattern)
            // stores them directly; any other keeps zero-and-accumulate: its
            // count is stored negated
            for (int64_t c = sl * kSlotSlice; c < c1; ++c) {
                const int k0 = cf_off[size_t(c)], k1 = cf_off[size_t(c) + 1];
                int covered = 0;
                bool uniq = true;
                for (int k = k0; k < k1 && uniq; ++k) {
                    const int e = cf_slot[size_t(k)];
                    if (e < 0)
                        continue;
                    if (ci[size_t(e)] == int(c))
                        uniq = false;
                    for (int k2 = k0; k2 < k && uniq; ++k2)
                        if (cf_slot[size_t(k2)] == e)
                            uniq = false;
                    ++covered;
                }
                if (!uniq || covered != rp[size_t(c) + 1] - rp[size_t(c)] - 1)
                    cnt[size_t(c)] = -cnt[size_t(c)];
            }
            soff[size_t(sl)] = int(tot);
            tot += int64_t(kSlotSlice) * w;
            if (tot > INT_MAX)
                return set_error(MI_ERR_BAD_ARG, "mesh too large for the sliced slot arrays");
        }
        soff[size_t(nsl)] = int(tot);
        std::vector<int> smap(size_t(std::max<int64_t>(tot, 1)), -1);
        for (int64_t c = 0; c < nC; ++c) {
            const int64_t q0 = soff[size_t(c / kSlotSlice)] + c % kSlotSlice;
            for (int j = 0; j < std::abs(cnt[size_t(c)]); ++j)
                smap[size_t(q0 + int64_t(kSlotSlice) * j)] = cf_off[size_t(c)] + j;
        }
        MI_TRY(up(M->ss_off, soff.data(), nsl + 1, s));
        MI_TRY(up(M->ss_cnt, cnt.data(), nC, s));
        DevBuf<int> d_smap;
        MI_TRY(d_smap.upload(smap.data(), smap.size(), s));
        const size_t m_ = smap.size();
        MI_TRY(M->ss_face.alloc(m_));
        MI_TRY(M->ss_slot.alloc(m_));
        MI_TRY(M->ss_oth.alloc(m_));
        MI_TRY(M->ss_tag.alloc(m_));
        MI_TRY(M->ss_s2.alloc(m_));
        MI_TRY(M->ss_ds.alloc(m_));
        if (tot > 0) {
            k_slice_slots<<<reduce_grid(tot, 4), kBlock, 0, s>>>(
                M->dev(), d_smap.p, tot, M->ss_face.p, M->ss_slot.p, M->ss_oth.p, M->ss_tag.p,
                M->ss_s2.p, M->ss_ds.p);
            MI_HIP(hipGetLastError());
        }
        MI_HIP(hipStreamSynchronize(s)); // d_smap is freed on leaving the scope
    }
    MI_HIP(hipStreamSynchronize(s));
    *out = M.release();
    return MI_OK;
}

} // namespace

extern "C" {

MI_API int mi_mesh_create(const mi_mesh_soa* m, const int32_t* bc_of_zone, int n_zones,
                          mi_mesh** out)
{
    return mesh_create(m, bc_of_zone, n_zones, nullptr, out);
}

MI_API int mi_mesh_create_twins(const mi_mesh_soa* m, const int32_t* bc_of_zone, int n_zones,
                                const int32_t* twin_faces, mi_mesh** out)
{
    if (!twin_faces)
        return set_error(MI_ERR_BAD_ARG, "null twin table");
    return mesh_create(m, bc_of_zone, n_zones, twin_faces, out);
}

MI_API int mi_simplec_denominators(mi_mesh* M, const double* aP, const mi_csr* A_mom,
                                   double* den)
{
    clear_error();
    if (!M || !A_mom || (M->n_cells > 0 && (!aP || !den)))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    if (A_mom->pat != M->pat)
        return set_error(MI_ERR_BAD_ARG, "the momentum matrix is not on the mesh's pattern");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    MI_TRY(launch_simplec_denominators(M, A_mom->val.p, aP, den, s));
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

MI_API int mi_mesh_info(const mi_mesh* M, int64_t* n_cells, int64_t* n_faces, int64_t* nnz)
{
    if (!M)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    if (n_cells)
        *n_cells = M->n_cells;
    if (n_faces)
        *n_faces = M->n_faces;
    if (nnz)
        *nnz = M->pat->nnz;
    return MI_OK;
}

MI_API int mi_mesh_create_matrix(mi_mesh* M, mi_csr** out)
{
    if (!M || !out)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    return make_csr_on_pattern(M->pat, out);
}

MI_API int mi_assemble_pressure(mi_mesh* M, double rho, const double* aP, const double* gamma,
                                const double* phi, mi_csr* A, double* rhs, int* pinned)
{
    if (!M || !A || !rhs)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    if (A->pat != M->pat)
        return set_error(MI_ERR_BAD_ARG, "matrix was not created on this mesh's pattern");
    MI_TRY(ensure_device());
    const int pin = (!M->any_dirichlet && M->n_cells > 0) ? 1 : 0;
    hipStream_t s = default_stream();
    MI_TRY(launch_assemble_pressure(M, rho, aP, gamma, phi, pin, A->val.p, rhs, s));
    MI_HIP(hipStreamSynchronize(s));
    A->values_changed(false);
    if (pinned)
        *pinned = pin;
    return MI_OK;
}

MI_API int mi_assemble_momentum(mi_mesh* M, int comp, double mu, double rho, const double* phi,
                                const double* pgradV, const double* dt, const double* Uold,
                                const double* Ucomp, double urf, const double* bc_value_host,
                                mi_csr* A, double* rhs)
{
    if (!M || !A || !rhs || !phi || comp < 0 || comp > 2 || !(urf > 0.0))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    if (A->pat != M->pat)
        return set_error(MI_ERR_BAD_ARG, "matrix was not created on this mesh's pattern");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    DevBuf<double> bv;
    MI_TRY(mesh_bc_values(M, bc_value_host, bv, s));
    MI_TRY(launch_assemble_momentum(M, comp, mu, rho, phi, pgradV, dt, Uold, nullptr, Ucomp, urf,
                                    bv.p, A->val.p, rhs, s));
    MI_HIP(hipStreamSynchronize(s));
    A->values_changed(false);
    return MI_OK;
}

MI_API int mi_face_gradient(mi_mesh* M, double rho, const double* aP, int pinned, const double* x,
                            const double* lambda, double* grad)
{
    if (!M || (M->n_faces && (!x || !lambda || !grad)))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    if (M->n_faces == 0)
        return MI_OK;
    MI_TRY(face_gradient_setup(M, s));
    const int n = int(M->n_cells);
    MI_TRY(M->fg_x.alloc(size_t(n)));
    MI_TRY(M->fg_l.alloc(size_t(n)));
    if (aP)
        MI_TRY(M->fg_a.alloc(size_t(n)));
    k_fg_gather<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, M->fg_perm.p, x, lambda, aP, M->fg_x.p,
                                                       M->fg_l.p, aP ? M->fg_a.p : nullptr);
    k_face_gradient<<<grid_for(M->n_faces, kBlock), kBlock, 0, s>>>(
        int(M->n_faces), M->fg_o.p, M->fg_n.p, M->fg_s2.p, M->fg_ds.p, rho,
        aP ? M->fg_a.p : nullptr, pinned ? M->fg_pin : -1, M->fg_x.p, M->fg_l.p, grad);
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

} // extern "C"

// ------------------------------------------------- allocation-free face flux
// mi_face_flux_noalloc keeps the zone velocity table, which mi_face_flux
// uploads into a buffer of its own on every call, in this file-scope device
// array instead.  Each GPU holds its own copy at its own address for the
// life of the process; host code never passes the array's name to a kernel,
// only the address hipGetSymbolAddress returns for the current device.
namespace
{
constexpr int kNoAllocMaxZones = 1024; // 24 KiB of zone velocities
constexpr int kZoneChunk = 128;        // zones per upload launch (3 KiB of arguments)
constexpr int kMaxDevices = 64;
} // namespace

static __device__ __align__(256) double scratch[kNoAllocMaxZones * 3];

namespace
{
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_table(double** out)
{
    int dev = 0;
    MI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(MI_ERR_DEVICE, "device %d beyond the scratch table of %d devices", dev,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(MI_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                             hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return MI_OK;
}

// Up to kZoneChunk zone velocities passed by value, so that filling the
// table needs neither a host staging copy nor a memset.
struct ZoneChunk
{
    int first;
    int count;
    double v[kZoneChunk * 3];
};

__global__ __launch_bounds__(kBlock) void k_store_zones(ZoneChunk c, double* __restrict__ table)
{
    for (int t = threadIdx.x; t < c.count * 3; t += kBlock)
        table[size_t(c.first) * 3 + size_t(t)] = c.v[t];
}
} // namespace

extern "C" {

MI_API int mi_face_flux_noalloc(mi_mesh* M, double rho, const double* Ux, const double* Uy,
                                const double* Uz, const double* bc_value_host, double* phi)
{
    if (!M || !Ux || !Uy || !Uz || !phi)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    if (M->n_zones > kNoAllocMaxZones)
        return set_error(MI_ERR_BAD_ARG,
                         "mi_face_flux_noalloc: %d boundary zones exceed its table of %d "
                         "(use mi_face_flux)",
                         M->n_zones, kNoAllocMaxZones);
    MI_TRY(ensure_device());
    double* table = nullptr;
    MI_TRY(scratch_table(&table));
    hipStream_t s = default_stream();
    // the zones this mesh can read (an inlet face's zone is < n_zones); rows
    // a previous call left beyond them are never read
    for (int z0 = 0; z0 < M->n_zones; z0 += kZoneChunk) {
        ZoneChunk c;
        c.first = z0;
        c.count = std::min(kZoneChunk, M->n_zones - z0);
        for (int t = 0; t < c.count * 3; ++t)
            c.v[t] = bc_value_host ? bc_value_host[size_t(z0) * 3 + size_t(t)] : 0.0;
        k_store_zones<<<1, kBlock, 0, s>>>(c, table);
        MI_HIP(hipGetLastError());
    }
    MI_TRY(launch_face_flux(M, rho, Ux, Uy, Uz, table, phi, s));
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

MI_API int mi_face_flux(mi_mesh* M, double rho, const double* Ux, const double* Uy,
                        const double* Uz, const double* bc_value_host, double* phi)
{
    if (!M || !Ux || !Uy || !Uz || !phi)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    DevBuf<double> bv;
    MI_TRY(mesh_bc_values(M, bc_value_host, bv, s));
    MI_TRY(launch_face_flux(M, rho, Ux, Uy, Uz, bv.p, phi, s));
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

MI_API int mi_mesh_destroy(mi_mesh* M)
{
    delete M;
    return MI_OK;
}

} // extern "C"
