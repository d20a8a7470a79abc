/*
 * simall_amd.h — C ABI of the MI355X (gfx950) pressure-correction solve backend.
 *
 * Plain C: pointers, sizes and integer status codes.  No STL or torch types
 * cross this boundary.  Each group of entry points names the reference
 * interface it replaces (paths relative to the SimAll source tree).
 *
 * Conventions
 *   - Every function returns an mi_status (MI_OK == 0).  On failure the
 *     thread-local mi_last_error() describes the cause.
 *   - "host" arguments are ordinary CPU memory, borrowed for the call.
 *     "device" arguments are HIP device pointers (mi_device_alloc, hipMalloc
 *     or a torch CUDA tensor's data_ptr()).
 *   - Handles (mi_csr, mi_solver, mi_mesh) own their device memory and are
 *     released by the matching *_destroy.  A handle is used by one host
 *     thread at a time; distinct handles may be used concurrently.
 *   - CSR: int32 rowPtr[n+1], int32 colIdx[nnz] (sorted per row), double
 *     values[nnz], as solver::CSRMatrix (src/solver/CSRMatrix.hpp:16-23).
 */
#ifndef SIMALL_AMD_H
#define SIMALL_AMD_H

#include <stddef.h>
#include <stdint.h>

#if defined(_WIN32)
#define MI_API __declspec(dllexport)
#else
#define MI_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define MI_ABI_VERSION 1

/* ------------------------------------------------------------- status codes
 * The reference's solvers never throw: non-convergence is the returned
 * iteration count plus last_residual (src/solver/LinearSolvers.cpp:451),
 * GPU failures are logged with default stats (src/gpu/CgGpu.cpp:105-116).
 * These codes carry the same distinctions. */
typedef enum {
    MI_OK = 0,
    MI_ERR_NOT_CONVERGED = 1, /* max iterations reached; x holds the last iterate */
    MI_ERR_BREAKDOWN = 2,     /* Krylov breakdown (p.Ap == 0, rho == 0, ...) */
    MI_ERR_BAD_ARG = 3,
    MI_ERR_DEVICE = 4,        /* HIP runtime error */
    MI_ERR_NO_DEVICE = 5,     /* no usable gfx950 device */
    MI_ERR_UNSUPPORTED = 6,   /* solver/preconditioner combination not on device */
    MI_ERR_ZERO_PIVOT = 7     /* ILU(0) met a zero pivot */
} mi_status;

/* Same ordinals as solver::LinearSolverKind / PreconditionerKind
 * (src/solver/Solver.hpp:67-81). */
enum { MI_SOLVER_GMRES = 0, MI_SOLVER_BICGSTAB = 1, MI_SOLVER_CG = 2, MI_SOLVER_TFQMR = 3 };
enum { MI_PC_JACOBI = 0, MI_PC_ILU = 1, MI_PC_SSOR = 2, MI_PC_AMG = 3, MI_PC_NONE = 4 };
/* Arithmetic of the Krylov iterations. */
enum { MI_PRECISION_FP64 = 0, MI_PRECISION_MIXED = 1 };

/* ---------------------------------------------------------------- runtime */
/* Replaces CudaContext (src/gpu/CudaContext.hpp:76-127) and
 * Device::is_available (src/gpu/Device.cpp:6-9). */
MI_API int mi_get_abi_version(void);
MI_API int mi_init(int device);
MI_API int mi_device_count(int* count);
MI_API const char* mi_last_error(void);
MI_API int mi_device_synchronize(void);

/* Device memory: DeviceMemoryPool::allocate/release
 * (src/gpu/DeviceMemoryPool.hpp:64-104) and HostDeviceMirror::to_device /
 * to_host (src/gpu/HostDeviceMirror.hpp:121-151). */
MI_API int mi_device_alloc(size_t bytes, void** ptr);
MI_API int mi_device_free(void* ptr);
MI_API int mi_memcpy_h2d(void* dst_device, const void* src_host, size_t bytes);
MI_API int mi_memcpy_d2h(void* dst_host, const void* src_device, size_t bytes);
MI_API int mi_memcpy_d2d(void* dst_device, const void* src_device, size_t bytes);
MI_API int mi_memset_device(void* dst_device, int value, size_t bytes);

/* ------------------------------------------------- vector and sparse kernels
 * The free functions of src/gpu/GpuKernels.hpp:47-108 on device pointers,
 * double precision, element counts.  `stream` is a hipStream_t or NULL (the
 * library's stream).  dot / nrm2 block until the result is on the host,
 * like the reference. */
MI_API int mi_axpy(double alpha, const void* x, void* y, size_t n, void* stream);
MI_API int mi_scal(double alpha, void* x, size_t n, void* stream);
MI_API int mi_copy(const void* src, void* dst, size_t n, void* stream);
MI_API int mi_dot(const void* x, const void* y, size_t n, double* result, void* stream);
MI_API int mi_nrm2(const void* x, size_t n, double* result, void* stream);
MI_API int mi_axpby(double a, const void* x, double b, void* y, size_t n, void* stream);
MI_API int mi_xpby(const void* x, double beta, void* y, size_t n, void* stream);
MI_API int mi_spmv_csr(size_t n, const void* rowPtr, const void* colIdx, const void* values,
                       const void* x, void* y, void* stream);
/* z[i] = r[i] / diag[i] (GpuKernels.hpp:85) */
MI_API int mi_jacobi_apply(const void* diag, const void* r, void* z, size_t n, void* stream);
/* In-place ILU(0) of device CSR values (GpuKernels.hpp:87-93; factorized on the
 * host like the reference, then written back).  *ok = 0 as the reference's
 * false return (GpuKernels.cpp:345-388): a row without a stored diagonal
 * (values left unchanged), or a zero pivot (the factorization stops there,
 * rows before it factored). */
MI_API int mi_ilu0_factor_inplace(size_t n, const void* rowPtr, const void* colIdx, void* values,
                                  int* ok);
/* z = U^-1 L^-1 r with level-scheduled device sweeps (GpuKernels.hpp:95-107).
 * A row without a stored diagonal, or with a zero one, is divided by 1. */
MI_API int mi_ilu0_apply(size_t n, const void* rowPtr, const void* colIdx, const void* values,
                         const void* r, void* z, void* stream);

/* ------------------------------------------------------------ CSR handles
 * Device-resident copy of a solver::CSRMatrix (src/solver/CSRMatrix.hpp:16-36);
 * the role of CgGpu::set_matrix's HostCsrView upload (src/gpu/CgGpu.hpp:78-96).
 * Arrays are copied; the caller keeps ownership of its host arrays. */
typedef struct mi_csr mi_csr;
MI_API int mi_csr_create(int64_t n, int64_t nnz, const int32_t* rowPtr, const int32_t* colIdx,
                         const double* values, mi_csr** out);
/* New values on the same pattern (the SIMPLE pattern is built once and
 * re-filled every outer iteration: src/solver/SimpleAlgorithm.cpp:49-98).
 * Solvers set up on this matrix rebuild their preconditioner on next solve. */
MI_API int mi_csr_update_values(mi_csr* A, const double* values_host);
/* Device-side variant.  values_device may be the handle's own value array
 * (mi_csr_values_device): then nothing is copied and the call only marks the
 * values changed, which a caller must do after writing through that pointer. */
MI_API int mi_csr_update_values_device(mi_csr* A, const double* values_device);
MI_API int mi_csr_get_values(const mi_csr* A, double* values_host);
MI_API int mi_csr_get_pattern(const mi_csr* A, int32_t* rowPtr_host, int32_t* colIdx_host);
MI_API int mi_csr_info(const mi_csr* A, int64_t* n, int64_t* nnz);
/* Device pointer to the FP64 values (for kernels that fill them in place).
 * The handle's derived copies (host mirror, SELL-64 and column-ordered value
 * streams, FP32 forms, cached preconditioners) are marked stale once, by
 * this call.  After every later write through the pointer, call
 * mi_csr_update_values_device(A, *values_device): with the handle's own
 * pointer it copies nothing and only marks the values changed.  Without it,
 * solves and SpMVs keep using the copies derived from the earlier values. */
MI_API int mi_csr_values_device(mi_csr* A, double** values_device);
/* y = A x with host arrays (CSRMatrix::spmv, src/solver/CSRMatrix.hpp:25-35). */
MI_API int mi_csr_spmv(mi_csr* A, const double* x_host, double* y_host);
/* y = A x on device.  precision MI_PRECISION_FP64: double x,y;
 * MI_PRECISION_MIXED: float values, float x,y (the inner-iteration form). */
MI_API int mi_csr_spmv_device(mi_csr* A, int precision, const void* x_device, void* y_device,
                              void* stream);
/* Times `reps` back-to-back SpMV launches on one stream with HIP events and
 * returns the average launch duration in microseconds. */
MI_API int mi_csr_time_spmv(mi_csr* A, int precision, const void* x_device, void* y_device,
                            int reps, double* avg_us);
/* The same with the matrix streams out of every cache: the launches rotate
 * over copies of the matrix (row pointers, columns, values) whose total
 * exceeds flush_bytes (at least 512 MiB: more than the 256 MB Infinity
 * Cache), so each launch finds its own copy evicted by the others.  A
 * back-to-back loop on one copy of a matrix that fits in that cache would
 * time cache bandwidth, not HBM.  Launches are queued back to back between
 * two events, as in mi_csr_time_spmv. */
MI_API int mi_csr_time_spmv_cold(mi_csr* A, int precision, const void* x_device, void* y_device,
                                 int reps, int64_t flush_bytes, double* avg_us);
/* Explicit transpose A^T as a new handle: the transposed pattern is built on
 * the host once, values are gathered on the device.  Replaces the scatter
 * product adjoint::spmv_transpose (src/adjoint/DiscreteAdjoint.cpp:13-22)
 * with a CSR of A^T that every SpMV-shaped kernel can stream.  At must not
 * outlive A; mi_csr_transpose_values refreshes At after A's values change. */
MI_API int mi_csr_transpose(const mi_csr* A, mi_csr** At);
MI_API int mi_csr_transpose_values(const mi_csr* A, mi_csr* At);
MI_API int mi_csr_destroy(mi_csr* A);

/* ---------------------------------------------------------------- solvers
 * Replaces ILinearSolver / make_linear_solver (src/solver/Solver.hpp:83-102,
 * src/solver/Solver.cpp:25-39) and the unwired CgGpu / BicgstabGpu
 * (src/gpu/CgGpu.hpp:49-125).  The first five fields are LinearSolverConfig
 * with the same defaults (tolerance 1e-8, 1000 iterations, restart 30). */
typedef struct {
    int kind;           /* MI_SOLVER_CG, _BICGSTAB or _GMRES on the device (TFQMR: unsupported) */
    int preconditioner; /* MI_PC_* (None and unknown kinds act as Jacobi, as
                           make_preconditioner: src/solver/LinearSolvers.cpp:240-254) */
    double tolerance;   /* stop when ||r||/max(||b||,1e-30) < tolerance; GMRES measures
                           the preconditioned residual ||M^-1 r|| as GMRESmSolver::solve
                           (src/solver/LinearSolvers.cpp:281-326) */
    int max_iterations;
    int restart;        /* GMRES cycle length m, max(1, restart), at most 256 on the device */
    int precision;      /* MI_PRECISION_FP64 or MI_PRECISION_MIXED (GMRES: FP32 Arnoldi
                           cycles, each restarted and tested on the FP64 residual) */
    int warm_start;     /* 0: x0 = 0 as the CPU solvers (LinearSolvers.cpp:424); 1: use x */
    int true_residual;  /* 0: stop on the recurrence residual (reference); 1: the FP64
                           true residual b - Ax must meet the tolerance (CG, BiCGStab;
                           GMRES cycles always restart from it) */
    double inner_tolerance; /* mixed: relative reduction of each FP32 inner solve (1e-4) */
    int max_outer;      /* mixed: refinement sweeps */
    int reorder;        /* -1 auto (n >= 100000: multicolour for ILU(0), else reverse
                           Cuthill-McKee), 0 none, 1 reverse Cuthill-McKee, 2 multicolour */
    int check_interval; /* iterations between host convergence polls (0 = auto) */
    /* smoothed-aggregation AMG */
    int amg_max_levels;
    int amg_coarse_size;
    double amg_strength;   /* aggregation strength threshold */
    double amg_jacobi_omega; /* prolongator smoothing weight (scaled by 1/rho(D^-1A)) */
    int amg_smoother;      /* 0 = Chebyshev, 1 = damped Jacobi */
    int amg_sweeps;        /* pre = post sweeps / Chebyshev degree */
    int amg_coarse_sweeps; /* the same on levels below the finest (default 1; 0: amg_sweeps) */
    int keep_preconditioner; /* 1: new values of the bound matrix keep the preconditioner
                                built on earlier values (the operator is always current) */
    int use_graph;           /* 1: replay captured HIP graphs (AMG V-cycles; the batched
                                solve's check_interval iterations), 0: launch each kernel */
    int amg_resetup;         /* 1 (default): when the bound matrix keeps the pattern of the
                                last full AMG setup, a setup for new values recomputes the
                                hierarchy's values on the device with that setup's aggregates,
                                strong couplings and sparsity patterns (lambda_max, smoothed P,
                                R = P^T, Galerkin R A P, coarsest inverse); 0: a full host
                                setup for every new set of values, as
                                AMGPreconditioner::setup (src/solver/AMGPreconditioner.cpp:248-326)
                                at every solve (src/solver/LinearSolvers.cpp:422) */
    double amg_truncation;   /* >= 0: each row of a smoothed prolongator drops the entries
                                below amg_truncation x its largest |entry| and scales the rest
                                to keep the row sum (default 0.1); 0 keeps every entry.  A
                                numeric re-setup keeps the set of entries of the full setup. */
} mi_solver_cfg;

/* GpuSolverStats (src/gpu/CgGpu.hpp:66-73) plus refinement counters. */
typedef struct {
    int iterations;        /* Krylov iterations (all inner iterations when mixed); GMRES
                              counts as the reference: the step that met the tolerance not */
    int outer_iterations;  /* refinement sweeps (mixed); GMRES restart cycles; 1 otherwise */
    int converged;
    int status;            /* mi_status of the solve */
    double initial_residual; /* ||b - A x0|| */
    double final_residual;   /* ||r|| at exit (recurrence or true, see cfg) */
    double relative_residual; /* final_residual / max(||b||,1e-30) == last_residual() */
    double true_relative_residual; /* FP64 ||b - A x|| / ||b|| computed on device at exit */
    double setup_seconds;
    double solve_seconds;
} mi_stats;

typedef struct mi_solver mi_solver;
MI_API void mi_solver_cfg_default(mi_solver_cfg* cfg);
MI_API int mi_solver_create(const mi_solver_cfg* cfg, mi_solver** out);
/* Binds A; builds the preconditioner now (the reference rebuilds it inside
 * every solve: src/solver/LinearSolvers.cpp:422; here it is rebuilt only when
 * A's values changed). */
MI_API int mi_solver_setup(mi_solver* s, mi_csr* A);
/* Host b, x (x is the initial guess only when warm_start). Returns MI_OK,
 * MI_ERR_NOT_CONVERGED or MI_ERR_BREAKDOWN; x and stats are written in all
 * three cases. */
MI_API int mi_solver_solve(mi_solver* s, const double* b_host, double* x_host, mi_stats* stats);
MI_API int mi_solver_solve_device(mi_solver* s, const double* b_device, double* x_device,
                                  mi_stats* stats);
/* IPreconditioner::apply (src/solver/LinearSolvers.hpp:28-35): z = M^-1 r
 * with the preconditioner mi_solver_setup built (FP32 copy when the solver's
 * precision is MI_PRECISION_MIXED).  Host arrays in the caller's numbering. */
MI_API int mi_solver_precond_apply(mi_solver* s, const double* r_host, double* z_host);
/* Preconditioner levels / operator complexity and the pattern bandwidth
 * before and after the solver's internal reordering (-1: not reordered). */
MI_API int mi_solver_info(const mi_solver* s, int* levels, double* operator_complexity,
                          int* bandwidth_before, int* bandwidth_after);
/* Times `reps` SpMVs with the solver's working matrix (after its internal
 * reordering) on the library stream with HIP events: average microseconds per
 * launch and the algorithmic bytes of one launch (values + columns + row
 * pointers + x once + y). */
MI_API int mi_solver_time_spmv(mi_solver* s, int precision, int reps, double* avg_us,
                               double* bytes);
/* mi_solver_time_spmv with mi_csr_time_spmv_cold's rotation over copies. */
MI_API int mi_solver_time_spmv_cold(mi_solver* s, int precision, int reps, int64_t flush_bytes,
                                    double* avg_us, double* bytes);
/* Adjoint solve A^T lambda = g (device vectors) with the solver's
 * configuration on the explicit transpose of the matrix bound by
 * mi_solver_setup; the transpose and its preconditioner are rebuilt when the
 * matrix's values change.  The adjoint-state solve of
 * adjoint::solve_discrete_adjoint (src/adjoint/DiscreteAdjoint.cpp:150-168),
 * which uses GMRES(30) on A^T there. */
MI_API int mi_adjoint_solve(mi_solver* s, const double* g_device, double* lambda_device,
                            mi_stats* stats);
/* ILinearSolver::last_residual (src/solver/Solver.hpp:99). */
MI_API double mi_solver_last_residual(const mi_solver* s);
/* Relative residual after each Krylov iteration of the last solve. */
MI_API int mi_solver_history(const mi_solver* s, double* out, int capacity, int* count);
MI_API int mi_solver_destroy(mi_solver* s);

/* ---------------------------------------------------- batched equations
 * n_eq (1..64) systems A_e x_e = b_e on one CSR pattern, solved together.
 * SimAll solves each scalar transport equation (energy, k/omega, species)
 * with its own ILinearSolver::solve on an identical pattern
 * (src/solver/ScalarTransport.cpp:84-242, the solve at :230-232), and the
 * three momentum components one after another on one matrix
 * (src/solver/SimpleAlgorithm.cpp:529-532, solve_momentum_component
 * :492-507); one mi_batch_solve replaces those n_eq calls.  Per equation the
 * iteration is PCG from x = 0 (CGSolver::solve,
 * src/solver/LinearSolvers.cpp:420-452), preconditioned BiCGStab
 * (BiCGSTABSolver::solve, :360-404) or left-preconditioned GMRES(m)
 * (GMRESmSolver::solve, :266-340; one cycle of at most m <= 32 Arnoldi steps
 * per refinement sweep, classical Gram-Schmidt twice) for nonsymmetric
 * (convective) equations, with Jacobi, ILU(0) (ILU0Preconditioner, :78-151) or
 * no preconditioner, in FP32 with FP64 dot products, norms and iterate,
 * refined on the FP64 true residual until ||b_e - A_e x_e|| <= tolerance *
 * ||b_e|| (for GMRES too, where the reference tests the preconditioned
 * residual); equations converge and drop out independently.  cfg (nullable:
 * defaults) supplies kind = CG, BiCGStab or GMRES (restart), preconditioner =
 * Jacobi, ILU or None, tolerance, max_iterations (per equation, all sweeps;
 * Arnoldi steps for GMRES), inner_tolerance, max_outer and check_interval.
 * Vectors are equation-major: b[e * n + i].  The rows per workgroup of the
 * batched SpMV are chosen by timing candidates at creation (mi_batch_tuning).
 *
 * flags of mi_batch_create_ex:
 *   MI_BATCH_SHARED_VALUES  one matrix for every equation (set by
 *                           mi_batch_set_matrix), streamed once per launch
 *                           for all right-hand sides;
 *   MI_BATCH_ILU0           order the rows by two colours so that the ILU(0)
 *                           factorization and its two triangular sweeps run
 *                           on the device, row-parallel; the pattern must be
 *                           structurally symmetric, hold every diagonal and
 *                           be two-colourable (hexahedral and quadrilateral
 *                           meshes), else MI_ERR_UNSUPPORTED.  The factors are
 *                           those of the reference's IKJ loop on that
 *                           ordering, computed in FP64 from the values at
 *                           the first ILU(0) solve after they change. */
typedef struct mi_batch mi_batch;
enum { MI_BATCH_SHARED_VALUES = 1, MI_BATCH_ILU0 = 2 };
MI_API int mi_batch_create(const mi_csr* pattern, int n_eq, mi_batch** out);
MI_API int mi_batch_create_ex(const mi_csr* pattern, int n_eq, int flags, mi_batch** out);
/* Values of equation eq from a matrix with the batch's pattern (not for a
 * batch with shared values). */
MI_API int mi_batch_set_equation(mi_batch* b, int eq, const mi_csr* A);
/* Values of every equation from one matrix with the batch's pattern: the
 * shared matrix, or each equation's copy. */
MI_API int mi_batch_set_matrix(mi_batch* b, const mi_csr* A);
/* All values, equation-major: values_host[e * nnz + k] (not for a batch with
 * shared values). */
MI_API int mi_batch_set_values(mi_batch* b, const double* values_host);
/* z_e = M_e^-1 r_e for every equation with the batch's preconditioner
 * (MI_PC_JACOBI, MI_PC_ILU or MI_PC_NONE), in the FP32 arithmetic of the
 * iterations: IPreconditioner::apply (src/solver/LinearSolvers.cpp:55-63,
 * 126-151) for the batch's matrices.  Equation-major device vectors in the
 * caller's numbering. */
MI_API int mi_batch_apply_preconditioner(mi_batch* b, int preconditioner, const double* r_device,
                                         double* z_device);
/* Returns MI_OK when every equation converged, else MI_ERR_NOT_CONVERGED;
 * iterations[e] and rel_residual[e] (FP64 true) are written in both cases. */
MI_API int mi_batch_solve(mi_batch* b, const mi_solver_cfg* cfg, const double* b_host,
                          double* x_host, int* iterations, double* rel_residual);
MI_API int mi_batch_solve_device(mi_batch* b, const mi_solver_cfg* cfg, const double* b_device,
                                 double* x_device, int* iterations, double* rel_residual);
/* Autotuner result: rows per workgroup, time of one batched SpMV (all
 * equations active) and its algorithmic bytes. */
MI_API int mi_batch_tuning(const mi_batch* b, int* rows_per_block, double* spmm_us,
                           double* spmm_bytes);
/* The shapes the autotuner picked: the cap on equations per thread of the
 * SpMM and ILU(0) sweeps (4, 2 or 1; a launch takes the widest of those that
 * divides its width), their workgroup count and row walk (banded: 1, the
 * workgroups of each XCD sweeping its eighth of the rows together; 0, one
 * contiguous range per workgroup); and the vector stages' workgroup count
 * (one per 256 rows, at most 2048) with the time of one vector stage (the CG
 * update, all equations). */
MI_API int mi_batch_tuning_shape(const mi_batch* b, int* slots_per_thread_max,
                                 int* spmm_workgroups, int* banded, int* vector_workgroups,
                                 double* vector_us);
/* Replaces the autotuner's SpMM shape (the three values mi_batch_tuning_shape
 * reports first): the shape fixes the order in which the dot products'
 * partial sums are added, so two batches given the same shape, matrix and
 * right-hand sides return the same bits, where the timed choice may differ
 * from batch to batch.  slots_per_thread_max 1, 2 or 4; spmm_workgroups a
 * multiple of 8 up to 4096; banded 0 or 1. */
MI_API int mi_batch_set_shape(mi_batch* b, int slots_per_thread_max, int spmm_workgroups,
                              int banded);
/* Cost of a grid-wide barrier on this device: microseconds per cooperative
 * grid.sync() in a kernel of `workgroups` x 256 threads (each thread updates
 * one element between barriers), and per kernel boundary for the same work
 * as back-to-back launches.  The batched solver separates the dependent
 * stages of an iteration by kernel boundaries (replayed from a HIP graph)
 * rather than by grid barriers inside one persistent kernel; this measures
 * that choice.  No reference counterpart. */
MI_API int mi_time_grid_barrier(int workgroups, int iterations, double* sync_us,
                                double* launch_us);
MI_API int mi_batch_destroy(mi_batch* b);

/* ------------------------------------------------- host-side inspection
 * The smoothed-aggregation hierarchy and the reverse Cuthill-McKee ordering
 * that mi_solver_setup builds, computed on the host for inspection and tests
 * (no device needed).  The reference's AMG (AMGPreconditioner::setup,
 * src/solver/AMGPreconditioner.cpp:248-326) keeps its levels private; these
 * entries have no reference counterpart. */
typedef struct mi_amg_host mi_amg_host;
MI_API int mi_amg_host_build(int64_t n, const int32_t* rowPtr, const int32_t* colIdx,
                             const double* values, const mi_solver_cfg* cfg, mi_amg_host** out);
MI_API int mi_amg_host_levels(const mi_amg_host* h, int* levels, double* operator_complexity,
                              double* grid_complexity);
/* which: 0 = A_l, 1 = P_l (prolongation from level l+1), 2 = R_l = P_l^T,
 * 3 = dense inverse of the coarsest operator (row-major, values only),
 * 4 = the aggregate of every row of A_l in level l+1's numbering (colIdx
 *     receives nrows entries, -1 for a row in no aggregate; ncols = the
 *     aggregate count),
 * 5 = per entry of A_l, 1 where it is a strong coupling kept by the filtered
 *     prolongator smoothing (colIdx receives nnz entries).
 * 4 and 5 exist on every level but the coarsest of a hierarchy built here
 * (not of one read from a checkpoint). */
MI_API int mi_amg_host_level_info(const mi_amg_host* h, int level, int which, int64_t* nrows,
                                  int64_t* ncols, int64_t* nnz, double* lambda_max);
MI_API int mi_amg_host_level_get(const mi_amg_host* h, int level, int which, int32_t* rowPtr,
                                 int32_t* colIdx, double* values);
MI_API int mi_amg_host_destroy(mi_amg_host* h);
/* A host copy of the hierarchy an AMG solver applies now: after a device
 * re-setup (mi_solver_cfg.amg_resetup) the values recomputed on the device,
 * and the dense coarsest inverse the device computed.  No reference
 * counterpart (the reference's levels are private). */
MI_API int mi_solver_amg_host(mi_solver* s, mi_amg_host** out);
/* perm[new] = old, reverse Cuthill-McKee of the symmetrised pattern. */
MI_API int mi_rcm_permutation(int64_t n, const int32_t* rowPtr, const int32_t* colIdx,
                              int32_t* perm, int32_t* bandwidth_after);
/* perm[new] = old, the multicolour ordering of the ILU(0) solvers (greedy
 * colouring in reverse Cuthill-McKee order, rows grouped by colour); *colors
 * receives the colour count.  No reference counterpart (the reference
 * factorizes in the caller's order). */
MI_API int mi_color_permutation(int64_t n, const int32_t* rowPtr, const int32_t* colIdx,
                                int32_t* perm, int32_t* colors);

/* ------------------------------------------------------ mesh and assembly
 * The SoA mesh of meshing::Mesh (src/meshing/MeshStorage.hpp:37-74): uint64
 * owner/neighbour with UINT64_MAX (kBoundaryCell) on boundary faces, FP64
 * area vectors and centroids, uint32 boundary zones, cell->face CSR with
 * int32 offsets and uint64 face ids.  Host arrays, copied at creation. */
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
} mi_mesh_soa;

/* Boundary treatment of a zone, grouping BCType (src/solver/Solver.hpp:105-123)
 * by what the SIMPLE face loops do with it:
 *   WALL            Wall, NoSlipWall: U = 0; zero-gradient p'
 *   PRESSURE        PressureInlet/Outlet: Dirichlet p' = 0; outflow for U
 *   SYMMETRY        Symmetry: no contribution
 *   VELOCITY_INLET  VelocityInlet, MovingWall: U = zone value
 *   OUTFLOW         every other type and zones without a BoundarySpec
 *   PERIODIC        Periodic: faces paired with a face of another periodic
 *                   zone couple their cells (an unpaired face acts as OUTFLOW) */
enum { MI_BC_WALL = 0, MI_BC_PRESSURE = 1, MI_BC_SYMMETRY = 2, MI_BC_VELOCITY_INLET = 3,
       MI_BC_OUTFLOW = 4, MI_BC_PERIODIC = 5 };

typedef struct mi_mesh mi_mesh;
/* Uploads the mesh and builds the CSR pattern of SimpleAlgorithm::build_sparsity
 * (src/solver/SimpleAlgorithm.cpp:49-98) with the twin columns of periodic
 * faces.  bc_of_zone[z] is the MI_BC_* class of zone z (zones >= n_zones:
 * OUTFLOW).  Periodic zones are paired as build_periodic_pairs
 * (:101-186) pairs them: in ascending zone id two by two (the order of a
 * BoundarySpec list given in zone order), each face of the first zone with
 * the nearest face of the second after the mean translation. */
MI_API int mi_mesh_create(const mi_mesh_soa* mesh, const int32_t* bc_of_zone, int n_zones,
                          mi_mesh** out);
/* mi_mesh_create with the caller's twin table instead of the pairing:
 * twin_faces[f] = the face paired with boundary face f, or -1
 * (SimpleAlgorithm::set_periodic_pairs, SimpleAlgorithm.hpp:84-92; e.g. the
 * rotational pairs of PeriodicPairing.hpp).  Entries of interior faces are
 * ignored. */
MI_API int mi_mesh_create_twins(const mi_mesh_soa* mesh, const int32_t* bc_of_zone, int n_zones,
                                const int32_t* twin_faces, mi_mesh** out);
/* SIMPLEC denominators den[c] = max(aP[c] - sum_{k != c} |A_mom[c][k]|, 1e-30)
 * (src/solver/SimpleAlgorithm.cpp:378-387, 400-409) from a momentum matrix on
 * the mesh's pattern; mi_assemble_pressure with aP = den is the SIMPLEC
 * assembly.  Device arrays of n_cells. */
MI_API int mi_simplec_denominators(mi_mesh* m, const double* aP_device, const mi_csr* A_mom,
                                   double* den_device);
MI_API int mi_mesh_info(const mi_mesh* m, int64_t* n_cells, int64_t* n_faces, int64_t* nnz);
/* A zero-valued matrix on the mesh's pattern (A_p_ / A_mom_ storage). */
MI_API int mi_mesh_create_matrix(mi_mesh* m, mi_csr** out);
/* SimpleAlgorithm::assemble_pressure_correction (src/solver/SimpleAlgorithm.cpp:358-447)
 * with its periodic branch (:411-419); SIMPLE denominators, or SIMPLEC with
 * aP = mi_simplec_denominators.  Device arrays: aP[n_cells] (nullable: aP_f = 1),
 * gamma[n_faces] (nullable: 1) scaling each face coefficient
 * rho*gamma_f*|S|^2/max(aP_f*|d.S|,1e-30), phi[n_faces] (nullable: rhs 0).
 * Writes A's values and rhs[n_cells]; with no pressure-zone face the system
 * is pinned (A_00 *= 1e6, rhs_0 = 0) and *pinned = 1. */
MI_API int mi_assemble_pressure(mi_mesh* m, double rho, const double* aP_device,
                                const double* gamma_device, const double* phi_device, mi_csr* A,
                                double* rhs_device, int* pinned);
/* SimpleAlgorithm::assemble_momentum (src/solver/SimpleAlgorithm.cpp:240-355),
 * first-order upwind + diffusion for velocity component comp.  Device arrays:
 * phi[n_faces]; pgradV[n_cells] (nullable) = grad(p)_comp * V subtracted from
 * the rhs; dt[n_cells] (nullable: steady) adds rho*V/dt implicit Euler with
 * Uold; Ucomp the current component for under-relaxation urf.  Host
 * bc_value[n_zones*3] holds inlet velocities. */
MI_API int mi_assemble_momentum(mi_mesh* m, int comp, double mu, double rho,
                                const double* phi_device, const double* pgradV_device,
                                const double* dt_device, const double* Uold_device,
                                const double* Ucomp_device, double urf,
                                const double* bc_value_host, mi_csr* A, double* rhs_device);
/* Per-face gradient of an objective J(x) through the pressure system of
 * mi_assemble_pressure with face scales gamma: grad[f] = dJ/dgamma_f =
 * -lambda^T (dA/dgamma_f) x, where A^T lambda = dJ/dx (mi_adjoint_solve).
 * An interior face adds gamma_f G_f (e_o - e_n)(e_o - e_n)^T to A, a
 * pressure-boundary face gamma_f G_f e_o e_o^T, a periodic face
 * gamma_f G_f e_o (e_o - e_t)^T (t: the twin's cell), G_f = rho |S_f|^2 /
 * max(aP_f |d_f . S_f|, 1e-30) (src/solver/SimpleAlgorithm.cpp:373-397,
 * 411-431); with `pinned` the pinned cell's diagonal carries the factor 1e6
 * (:441-446).  Other faces get 0.  The gradient step
 * dJ/dalpha - (dR/dalpha)^T psi of src/adjoint/DiscreteAdjoint.cpp:170-182.
 * Device arrays: aP[n_cells] (nullable), x, lambda [n_cells], grad[n_faces]. */
MI_API int mi_face_gradient(mi_mesh* m, double rho, const double* aP_device, int pinned,
                            const double* x_device, const double* lambda_device,
                            double* grad_device);
/* SimpleAlgorithm::compute_face_fluxes (src/solver/SimpleAlgorithm.cpp:201-237),
 * periodic faces averaging with the twin's cell. */
MI_API int mi_face_flux(mi_mesh* m, double rho, const double* Ux_device, const double* Uy_device,
                        const double* Uz_device, const double* bc_value_host, double* phi_device);
/* mi_face_flux that allocates nothing, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight outer-iteration loops).
 * No counterpart in the reference.  The zone velocity table lives in a
 * library-owned device array of 1024 zones (24 KiB per GPU, held until the
 * process ends); a mesh with more zones is refused with MI_ERR_BAD_ARG.
 * Every call on a GPU shares that GPU's table: calls that overlap on one GPU
 * (several host threads) race. */
MI_API int mi_face_flux_noalloc(mi_mesh* m, double rho, const double* Ux_device,
                                const double* Uy_device, const double* Uz_device,
                                const double* bc_value_host, double* phi_device);
MI_API int mi_mesh_destroy(mi_mesh* m);

/* ----------------------------------------------------- gradient, correction
 * LeastSquaresGradient::evaluate (src/solver/Gradient.cpp:93-136) of a cell
 * field: device phi[n_cells] -> gx, gy, gz [n_cells].  The inverted moment
 * matrices (Gradient.cpp:42-91, built from the mesh geometry) are made on
 * first use and kept with the mesh.  Bit-exact to the reference. */
MI_API int mi_ls_gradient(mi_mesh* m, const double* phi_device, double* gx_device,
                          double* gy_device, double* gz_device);
/* SimpleAlgorithm::correct_fields (src/solver/SimpleAlgorithm.cpp:450-489):
 * p += urfP p'; U -= V / max(aP, 1e-30) grad(p') with the least-squares
 * gradient; on interior faces phi += rho |S|^2 / max(aP_f |d.S|, 1e-30) *
 * (p'_owner - p'_neighbour), aP_f = (aP_owner + aP_neighbour) / 2.  Device
 * arrays of n_cells (phi: n_faces, nullable to skip the flux correction). */
MI_API int mi_correct_fields(mi_mesh* m, double urfP, double rho, const double* aP_device,
                             const double* pprime_device, double* p_device, double* Ux_device,
                             double* Uy_device, double* Uz_device, double* phi_device);

/* ------------------------------------------------------- SIMPLE / PISO step
 * A device-resident SimpleAlgorithm (src/solver/SimpleAlgorithm.cpp:521-575)
 * or PisoAlgorithm (src/solver/PisoAlgorithm.cpp:27-70) on a mi_mesh: U, p,
 * the face flux phi, the momentum diagonal aP and both matrices stay in HBM
 * between iterations.  One iteration is: face fluxes; per velocity component
 * momentum assembly (least-squares pressure-gradient source, time term,
 * under-relaxation), solve and residual ||A U - b||; face fluxes; pressure
 * assembly, solve from p' = 0, ||rhs|| and correct_fields (PISO: n_correctors
 * times (n_nonorth + 1) pressure passes after the predictor).  Solves start
 * from x = 0 as the reference's CPU solvers do.  Fields start at zero; write
 * initial values through the pointers of mi_flow_fields. */
enum { MI_FLOW_SIMPLE = 0, MI_FLOW_PISO = 1,
       MI_FLOW_SIMPLEC = 2 /* SimplecAlgorithm (src/solver/SimplecAlgorithm.hpp): SIMPLE with
                              the denominators aP - sum |a_nb| in the pressure assembly */ };
enum { MI_TIME_EULER = 0, MI_TIME_BDF2 = 1 }; /* TemporalScheme ImplicitEuler, BDF2 */
typedef struct {
    int algorithm;       /* MI_FLOW_SIMPLE, MI_FLOW_PISO or MI_FLOW_SIMPLEC */
    double rho, mu;      /* density, dynamic viscosity (1, 1e-3) */
    double urfU, urfP;   /* under-relaxation: SIMPLE 0.7, 0.3; PISO 1, 1; SIMPLEC 0.9, 1 */
    double dt;           /* > 0: transient term; SIMPLE 0 (steady), PISO 1e-3 */
    int time_scheme;     /* MI_TIME_EULER (SIMPLE) or MI_TIME_BDF2 (PISO); BDF2 from
                            the second time step on, implicit Euler before */
    int n_correctors;    /* PISO pressure correctors (2) */
    int n_nonorth;       /* PISO extra pressure passes per corrector (0) */
    mi_solver_cfg momentum; /* momentum solves: GMRES(30) + ILU(0), 1e-8 (LinearSolverConfig) */
    mi_solver_cfg pressure; /* pressure-correction solves: CG + SA-AMG, 1e-8 */
    int momentum_batch;  /* 1: the three momentum components (one matrix, three right-hand
                            sides: SimpleAlgorithm.cpp:529-532) in one mi_batch solve on one
                            shared matrix, BiCGStab with ILU(0) (whatever momentum.kind: the
                            fastest batched kind here) and momentum.tolerance / max_iterations /
                            inner_tolerance / max_outer; 0 (default): three solves with
                            `momentum` */
} mi_flow_cfg;
/* SimpleResiduals (src/solver/SimpleAlgorithm.hpp:59-63). */
typedef struct {
    double mom[3];
    double cont;
} mi_flow_residuals;
typedef struct mi_flow mi_flow;
/* Defaults of SimpleOptions / PisoOptions for the algorithm. */
MI_API void mi_flow_cfg_default(int algorithm, mi_flow_cfg* cfg);
/* bc_value_host[n_zones * 3]: zone velocities (inlets), nullable. */
MI_API int mi_flow_create(mi_mesh* m, const double* bc_value_host, const mi_flow_cfg* cfg,
                          mi_flow** out);
/* Device pointers of the resident fields (each output nullable). */
MI_API int mi_flow_fields(mi_flow* f, double** Ux, double** Uy, double** Uz, double** p,
                          double** phi);
/* One outer iteration: SimpleAlgorithm::iterate / PisoAlgorithm::iterate. */
MI_API int mi_flow_iterate(mi_flow* f, mi_flow_residuals* res);
/* SimpleAlgorithm::advance_time_step: roll U^n (and U^{n-1}), n_inner iterations. */
MI_API int mi_flow_advance_time_step(mi_flow* f, int n_inner, mi_flow_residuals* res);
MI_API int mi_flow_destroy(mi_flow* f);

/* -------------------------------------------------------------- checkpoint
 * The SIMALLCP file of Checkpoint::write / read (src/io/Checkpoint.hpp:27-57,
 * src/io/Checkpoint.cpp:67-230): a header (magic, version 1, flags,
 * timestamp), CheckpointMeta, the mesh section and the field section; files
 * written here are read by the reference and the reverse.  The mesh section
 * needs the node data meshing::Mesh keeps besides mi_mesh_soa: */
typedef struct {
    int64_t n_nodes;
    const double* x;
    const double* y;
    const double* z;
    const int32_t* faceNodeOffsets;  /* [n_faces + 1] */
    const uint64_t* faceNodeIndices; /* [faceNodeOffsets[n_faces]] */
} mi_mesh_nodes;
/* CheckpointMeta (Checkpoint.hpp:35-40). */
typedef struct {
    int32_t iteration;
    double time;
    uint64_t settings_hash;
} mi_checkpoint_meta;
typedef struct {
    const char* name;
    const double* data; /* [n_cells], host */
} mi_named_field;
typedef struct {
    const char* name;
    const double* x; /* [n_cells] each, host */
    const double* y;
    const double* z;
} mi_named_vector;
/* Writes <path>.tmp and renames it over <path> (host arrays). */
MI_API int mi_checkpoint_write(const char* path, const mi_mesh_soa* mesh,
                               const mi_mesh_nodes* nodes, int n_scalars,
                               const mi_named_field* scalars, int n_vectors,
                               const mi_named_vector* vectors, const mi_checkpoint_meta* meta);
/* Reads a checkpoint into host memory: MI_ERR_BAD_ARG on a bad magic or a
 * truncated file, MI_ERR_UNSUPPORTED on a newer version. */
typedef struct mi_checkpoint mi_checkpoint;
MI_API int mi_checkpoint_open(const char* path, mi_checkpoint** out);
MI_API int mi_checkpoint_info(const mi_checkpoint* c, mi_checkpoint_meta* meta, int64_t* n_cells,
                              int64_t* n_faces, int64_t* n_nodes);
/* Copies up to capacity values of a named field; count receives its length. */
MI_API int mi_checkpoint_scalar(const mi_checkpoint* c, const char* name, double* out,
                                int64_t capacity, int64_t* count);
MI_API int mi_checkpoint_vector(const mi_checkpoint* c, const char* name, double* x, double* y,
                                double* z, int64_t capacity, int64_t* count);
MI_API int mi_checkpoint_close(mi_checkpoint* c);
/* Optional AMG trailer (no counterpart in the reference's format; the
 * reference's reader stops after the field section and ignores it): the SA-AMG
 * hierarchy appended after the fields, so that a restarted run that keeps its
 * preconditioner (mi_solver_cfg.keep_preconditioner) needs no AMG setup and
 * continues bit for bit.  levels = 0 when the file has none. */
MI_API int mi_checkpoint_amg_levels(const mi_checkpoint* c, int* levels);
/* Appends a host hierarchy (mi_amg_host_build) to a checkpoint file. */
MI_API int mi_checkpoint_append_amg(const char* path, const mi_amg_host* h);
/* The device-resident flow state as fields "p" and "U" (the FieldRegistry
 * names of SimpleAlgorithm), with the caller's mesh arrays.  flags
 * MI_CHECKPOINT_AMG also appends the pressure solver's AMG hierarchy
 * (MI_ERR_UNSUPPORTED if it has built none). */
enum { MI_CHECKPOINT_AMG = 1 };
MI_API int mi_flow_checkpoint_write(mi_flow* f, const char* path, const mi_mesh_soa* mesh,
                                    const mi_mesh_nodes* nodes, const mi_checkpoint_meta* meta,
                                    int flags);
/* Loads "p" and "U" from a checkpoint of a mesh with this flow's cell count
 * and clears the time history (the next time step is implicit Euler, as the
 * first step of a new flow).  An AMG trailer is adopted by the pressure
 * solver when it uses AMG with keep_preconditioner (its next solve then
 * skips the setup), and ignored otherwise. */
MI_API int mi_flow_checkpoint_read(mi_flow* f, const char* path, mi_checkpoint_meta* meta);

#ifdef __cplusplus
}
#endif

#endif /* SIMALL_AMD_H */
