/*
 * oracle.h — CPU restatement of the SimAll pressure-correction solve path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP backend:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product (libsimall_amd.so) never links, loads or calls it.
 *
 * Every function restates one reference routine loop-for-loop, including the
 * order of floating-point accumulation, so that results (iteration counts,
 * residual histories, assembled coefficients) match the reference's serial
 * code.  The reference file:line each function follows is cited beside it.
 *
 * Conventions: CSR with int32 rowPtr[n+1] / colIdx[nnz] (columns sorted per
 * row, diagonal present) and double values[nnz], as solver::CSRMatrix
 * (src/solver/CSRMatrix.hpp:16-23).
 *
 * Parity pins (which check fixes each function) are listed in the module
 * docstring of oracle/__init__.py; functions with no pin say so there.
 */
#ifndef SIMALL_ORACLE_H
#define SIMALL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ RNG */
/* SplitMix64 (Steele, Lea, Flood 2014). uniform = (x >> 11) * 2^-53. */
uint64_t orc_splitmix64(uint64_t* state);
double orc_uniform(uint64_t* state);
/* Box-Muller, cosine branch, one normal per two uniforms. */
double orc_normal(uint64_t* state);

/* ------------------------------------------------------------------ CSR */
typedef struct {
    int n;
    const int* rowPtr;
    const int* colIdx;
    const double* values;
} orc_csr;

/* y = A x, row-sequential (src/solver/CSRMatrix.hpp:25-35). */
void orc_spmv(const orc_csr* A, const double* x, double* y);
/* y = A^T x by scatter (src/adjoint/DiscreteAdjoint.cpp:13-22). */
void orc_spmv_transpose(const orc_csr* A, const double* x, double* y);

/* BLAS-1 (src/solver/LinearSolvers.cpp:25-48). Left-to-right sums. */
double orc_dot(const double* a, const double* b, int64_t n);
double orc_norm2(const double* a, int64_t n);
void orc_axpy(double alpha, const double* x, double* y, int64_t n);
void orc_scal(double a, double* x, int64_t n);

/* ------------------------------------------------------- preconditioners */
enum { ORC_PC_JACOBI = 0, ORC_PC_ILU = 1, ORC_PC_SSOR = 2, ORC_PC_AMG = 3, ORC_PC_NONE = 4 };

/* invDiag = 1/a_ii, 1 when the diagonal is missing or zero
 * (src/solver/LinearSolvers.cpp:54-65). */
void orc_jacobi_setup(const orc_csr* A, double* invDiag);

/* In-place ILU(0) on A's pattern, IKJ order, zero/missing pivots skipped
 * (src/solver/LinearSolvers.cpp:81-126). lu[nnz] receives L (unit, strict
 * lower) and U; diagPtr[n] the position of each diagonal (-1 if missing). */
void orc_ilu0_setup(const orc_csr* A, double* lu, int* diagPtr);
/* z = U^-1 L^-1 r (src/solver/LinearSolvers.cpp:127-151). */
void orc_ilu0_apply(const orc_csr* A, const double* lu, const int* diagPtr,
                    const double* r, double* z);

/* SSOR(omega) as the reference's SsorPreconditioner
 * (src/solver/LinearSolvers.cpp:163-236). */
void orc_ssor_apply(const orc_csr* A, const int* diagPtr, double omega,
                    const double* r, double* z);

/* ---------------------------------------------------------------- Krylov */
/* All solvers start from x = 0 (the reference ignores the input x) and stop
 * on the recurrence residual ||r||/max(||b||,1e-30) < tol.  They return the
 * iteration count (maxIter when not converged) and write last_residual.
 * history (nullable, size maxIter+1) receives the relative residual after
 * each iteration; *n_history its length. */
int orc_cg_solve(const orc_csr* A, int pc, double tol, int maxIter,
                 const double* b, double* x, double* last_residual,
                 double* history, int* n_history);
int orc_bicgstab_solve(const orc_csr* A, int pc, double tol, int maxIter,
                       const double* b, double* x, double* last_residual,
                       double* history, int* n_history);
int orc_tfqmr_solve(const orc_csr* A, int pc, double tol, int maxIter, const double* b,
                    double* x, double* last_residual);
int orc_gmres_solve(const orc_csr* A, int pc, double tol, int maxIter, int restart,
                    const double* b, double* x, double* last_residual);

/* Restarted GMRES on A^T with left-Jacobi (src/adjoint/DiscreteAdjoint.cpp:54-146).
 * Returns 1 when converged; iters / res receive the counters. */
int orc_gmres_transpose(const orc_csr* A, const double* b, double* x, int maxIter,
                        int restart, double tol, int* iters, double* res);

/* --------------------------------------------------------------- assembly */
/* Mesh arrays follow meshing::FaceStorage / CellStorage
 * (src/meshing/MeshStorage.hpp:37-74): owner/neighbour are uint64 with
 * UINT64_MAX marking a boundary face; cell->face CSR has int32 offsets and
 * uint64 face indices. */
typedef struct {
    int64_t n_cells;
    int64_t n_faces;
    const uint64_t* owner;
    const uint64_t* neighbor;
    const double* areaX;
    const double* areaY;
    const double* areaZ;
    const double* faceCx;
    const double* faceCy;
    const double* faceCz;
    const uint32_t* boundaryZone;
    const double* cellCx;
    const double* cellCy;
    const double* cellCz;
    const double* volume;
    const int32_t* cellFaceOffsets;
    const uint64_t* cellFaceIndices;
} orc_mesh;

/* Boundary treatment per zone id, as the reference's BCType families.  */
enum {
    ORC_BC_WALL = 0,             /* Wall / NoSlipWall: zero-gradient on p', fixed U=0 */
    ORC_BC_PRESSURE = 1,         /* PressureInlet/Outlet: Dirichlet p'=0; outflow for U */
    ORC_BC_SYMMETRY = 2,
    ORC_BC_VELOCITY_INLET = 3,   /* fixed U = value */
    ORC_BC_OUTFLOW = 4,
    ORC_BC_PERIODIC = 5          /* Periodic: paired with a face of another periodic zone */
};

/* Periodic face pairs of SimpleAlgorithm::build_periodic_pairs
 * (src/solver/SimpleAlgorithm.cpp:101-186): twin[f] = paired face or -1. */
void orc_periodic_twins(const orc_mesh* m, const int* bcOfZone, int nZones, int* twin);
/* SIMPLEC denominators max(aP_c - sum_{k != c} |a_ck|, 1e-30) of the momentum
 * matrix (src/solver/SimpleAlgorithm.cpp:378-387); the pressure assembly with
 * aP = den is the SIMPLEC assembly. */
void orc_simplec_denominators(const orc_csr* amom, const double* aP, double* den);
/* The functions below with a periodic twin table (NULL: none); the names
 * without _p pass NULL. */
int64_t orc_build_sparsity_p(const orc_mesh* m, const int* twin, int* rowPtr, int* colIdx);
int orc_assemble_pressure_p(const orc_mesh* m, const int* bcOfZone, int nZones, const int* twin,
                            double rho, const double* aP, const double* gamma,
                            const double* phi, const orc_csr* pattern,
                            double* values, double* rhs);
void orc_face_gradient_p(const orc_mesh* m, const int* bcOfZone, int nZones, const int* twin,
                         double rho, const double* aP, int pinned, const double* x,
                         const double* lam, double* grad);
void orc_assemble_momentum_p(const orc_mesh* m, const int* bcOfZone, int nZones, const int* twin,
                             const double* bcValue, int comp, double mu, double rho,
                             const double* phi, const double* pgradV, const double* dtCell,
                             const double* Uold, const double* Unm1, const double* Ucomp,
                             double urfU, const orc_csr* pattern, double* values, double* rhs);
void orc_face_flux_p(const orc_mesh* m, const int* bcOfZone, int nZones, const int* twin,
                     const double* bcValue, double rho, const double* Ux, const double* Uy,
                     const double* Uz, double* phi);

/* CSR pattern: row c = {c} U face neighbours, sorted, unique
 * (src/solver/SimpleAlgorithm.cpp:49-98, no periodic twins).
 * Pass rowPtr[n+1]; colIdx NULL to query nnz (returned). */
int64_t orc_build_sparsity(const orc_mesh* m, int* rowPtr, int* colIdx);

/* Pressure-correction assembly (src/solver/SimpleAlgorithm.cpp:358-447),
 * SIMPLE variant.  coef_f = rho*gamma_f*|S|^2 / max(aPf*|d.S|, 1e-30) with
 * aPf = 0.5*(max(aP_o,1e-30)+max(aP_n,1e-30)) (aP NULL: aPf = 1) and
 * gamma_f = 1 when gamma is NULL.  Pressure-zone faces add coef to A_oo;
 * walls add nothing.  rhs: b_o -= phi_f, b_n += phi_f (phi NULL: 0).
 * With no Dirichlet face: A_00 *= 1e6, b_0 = 0. Returns 1 if pinned. */
int orc_assemble_pressure(const orc_mesh* m, const int* bcOfZone, int nZones,
                          double rho, const double* aP, const double* gamma,
                          const double* phi, const orc_csr* pattern,
                          double* values, double* rhs);
/* grad[f] = dJ/dgamma_f = -lambda^T (dA/dgamma_f) x for the matrix of
 * orc_assemble_pressure, where A^T lambda = dJ/dx.  pinned: the value
 * orc_assemble_pressure returned. */
void orc_face_gradient(const orc_mesh* m, const int* bcOfZone, int nZones, double rho,
                       const double* aP, int pinned, const double* x, const double* lam,
                       double* grad);

/* Momentum assembly, first-order upwind + diffusion, one component
 * (src/solver/SimpleAlgorithm.cpp:240-355).  pgradV (nullable) is the
 * pressure-gradient source dp/dx_comp * V subtracted from the rhs.  dtCell
 * (nullable) adds the time term: implicit Euler rho*V/dt_c on the diagonal and
 * rho*V/dt_c*Uold_c on the rhs, or with Unm1 BDF2 (1.5 rho V/dt;
 * (2 Uold - 0.5 Unm1) rho V/dt); urfU then scales the diagonal as Patankar
 * (rhs += (a/urf - a)*Ucomp).  bcValue[zone*3+comp] gives the fixed velocity
 * of inlet zones. */
void orc_assemble_momentum(const orc_mesh* m, const int* bcOfZone, int nZones,
                           const double* bcValue, int comp, double mu, double rho,
                           const double* phi, const double* pgradV, const double* dtCell,
                           const double* Uold, const double* Unm1, const double* Ucomp,
                           double urfU, const orc_csr* pattern, double* values, double* rhs);

/* Least-squares cell gradient (src/solver/Gradient.cpp:42-136). */
void orc_ls_gradient(const orc_mesh* m, const double* phi, double* gx, double* gy, double* gz);
/* Field correction after the pressure solve (src/solver/SimpleAlgorithm.cpp:450-489):
 * p += urfP p'; U -= V/max(aP,1e-30) grad p'; interior-face flux correction
 * (phi nullable: skipped). */
void orc_correct_fields(const orc_mesh* m, double urfP, double rho, const double* aP,
                        const double* pp, double* p, double* Ux, double* Uy, double* Uz,
                        double* phi);

/* Face mass flux phi_f = rho * 0.5*(U_o+U_n) . S_f on interior faces
 * (src/solver/SimpleAlgorithm.cpp:201-237). */
void orc_face_flux(const orc_mesh* m, const int* bcOfZone, int nZones,
                   const double* bcValue, double rho, const double* Ux,
                   const double* Uy, const double* Uz, double* phi);

#ifdef __cplusplus
}
#endif

#endif
