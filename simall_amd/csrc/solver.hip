// solver.hip — mi_solver: device-resident CG with FP64 or mixed FP32/FP64
// arithmetic behind the ILinearSolver contract (src/solver/Solver.hpp:92-100).
//
// FP64 path: CGSolver::solve (src/solver/LinearSolvers.cpp:420-452) with the
// iteration kept on the device.  Mixed path: iterative refinement — the FP64
// true residual r = b - A x is formed on the device, an FP32 preconditioned
// CG solves A d = r to `inner_tolerance`, x += d in FP64, repeat until the
// FP64 true residual meets `tolerance`.  GMRES(m) (LinearSolvers.cpp:266-340)
// restarts every cycle from the FP64 residual; its mixed form runs the
// Arnoldi steps in FP32.
#include "solver.hpp"
#include "stream.hpp"
#include "amg.hpp"
#include "checkpoint.hpp"
#include "gmres.hpp"
#include "ilu0.hpp"
#include "reorder.hpp"

#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>

using namespace mi;

namespace mi
{

// ----------------------------------------------------------------- Jacobi
template <class T>
__global__ __launch_bounds__(kBlock) void k_scale(int n, const T* __restrict__ a,
                                                  const T* __restrict__ r, T* __restrict__ z)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n)
        z[i] = a[i] * r[i];
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int n, const double* __restrict__ a,
                                                   float* __restrict__ b)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n)
        b[i] = float(a[i]);
}

__global__ __launch_bounds__(kBlock) void k_to_f64(int n, const float* __restrict__ a,
                                                   double* __restrict__ b)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n)
        b[i] = double(a[i]);
}

namespace
{
struct JacobiPrecond final : Precond
{
    DevBuf<double> inv64;
    DevBuf<float> inv32;
    bool has64 = false, has32 = false;
    int setup(mi_csr* A, bool need64, bool need32, hipStream_t s) override
    {
        const Pattern& p = *A->pat;
        has64 = has32 = false;
        if (need64) {
            MI_TRY(inv64.alloc(size_t(p.n)));
            MI_TRY(launch_inv_diag(p, A->val.p, inv64.p, s));
            has64 = true;
        }
        if (need32) {
            MI_TRY(inv32.alloc(size_t(p.n)));
            MI_TRY(launch_inv_diag32(p, A->val.p, inv32.p, s));
            has32 = true;
        }
        return MI_OK;
    }
    const double* inv_diag64() const override { return has64 ? inv64.p : nullptr; }
    const float* inv_diag32() const override { return has32 ? inv32.p : nullptr; }
    int apply64(const double* r, double* z, hipStream_t s) override
    {
        if (!has64)
            return set_error(MI_ERR_UNSUPPORTED, "no FP64 inverse diagonal");
        const int n = int(inv64.n);
        k_scale<double><<<grid_for(n, kBlock), kBlock, 0, s>>>(n, inv64.p, r, z);
        MI_HIP(hipGetLastError());
        return MI_OK;
    }
    int apply32(const float* r, float* z, hipStream_t s) override
    {
        if (!has32)
            return set_error(MI_ERR_UNSUPPORTED, "no FP32 inverse diagonal");
        const int n = int(inv32.n);
        k_scale<float><<<grid_for(n, kBlock), kBlock, 0, s>>>(n, inv32.p, r, z);
        MI_HIP(hipGetLastError());
        return MI_OK;
    }
    const char* name() const override { return "jacobi"; }
};
} // namespace

std::unique_ptr<Precond> make_jacobi() { return std::make_unique<JacobiPrecond>(); }

// r = b - A x in FP64 (stored as r64 and/or rounded to r32), out = ||r||
struct EpiResidual
{
    const double* b;
    double* r64;
    float* r32;
    double acc = 0.0;
    double bi = 0.0;
    __device__ __forceinline__ void pre(int i) { bi = b[i]; }
    __device__ __forceinline__ void row(int i, double s)
    {
        const double ri = bi - s;
        if (r64)
            r64[i] = ri;
        if (r32)
            r32[i] = float(ri);
        acc += ri * ri;
    }
};

__global__ __launch_bounds__(kBlock) void k_true_residual(StreamMat<double> A,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ b,
                                                          double* __restrict__ r64,
                                                          float* __restrict__ r32,
                                                          double* partials)
{
    EpiResidual e{b, r64, r32};
    stream_rows(A, x, e);
    double v[1] = {e.acc};
    block_partials<1>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_norm(int n, const double* __restrict__ b,
                                                 double* partials)
{
    const RowRange rr = block_rows(n, gridDim.x, 1);
    double v[1] = {0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock)
        v[0] += b[i] * b[i];
    block_partials<1>(v, partials);
}

// dst[i] = src[perm[i]]
__global__ __launch_bounds__(kBlock) void k_gather(int n, const int* __restrict__ perm,
                                                   const double* __restrict__ src,
                                                   double* __restrict__ dst)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n)
        dst[i] = src[perm[i]];
}

// dst[perm[i]] = src[i]
__global__ __launch_bounds__(kBlock) void k_scatter(int n, const int* __restrict__ perm,
                                                    const double* __restrict__ src,
                                                    double* __restrict__ dst)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n)
        dst[perm[i]] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_gather_values(int64_t nnz, const int* __restrict__ vmap,
                                                          const double* __restrict__ src,
                                                          double* __restrict__ dst)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t k = int64_t(blockIdx.x) * kBlock + threadIdx.x; k < nnz; k += stride)
        dst[k] = src[vmap[k]];
}

} // namespace mi

mi_solver::~mi_solver()
{
    if (h_st)
        (void)hipHostFree(h_st);
    if (h_scal)
        (void)hipHostFree(h_scal);
    if (stream)
        (void)hipStreamDestroy(stream);
}

namespace
{

using Clock = std::chrono::steady_clock;

int true_residual(mi_solver* S, const double* b, const double* x, double* r64, float* r32,
                  double* out_dev)
{
    StreamMat<double> A;
    MI_TRY(stream_view64(S->M, S->stream, A));
    const unsigned G = stream_grid(A.nrb);
    k_true_residual<<<G, kBlock, 0, S->stream>>>(A, x, b, r64, r32, S->partials.p);
    MI_HIP(hipGetLastError());
    return launch_fin_sum(S->partials.p, G, true, out_dev, S->stream);
}

int norm2(mi_solver* S, int n, const double* v, double* out_dev)
{
    const unsigned G = range_grid(n, kBlock);
    k_norm<<<G, kBlock, 0, S->stream>>>(n, v, S->partials.p);
    MI_HIP(hipGetLastError());
    return launch_fin_sum(S->partials.p, G, true, out_dev, S->stream);
}

int read_scalars(mi_solver* S, int count)
{
    MI_HIP(hipMemcpyAsync(S->h_scal, S->dscal.p, sizeof(double) * size_t(count),
                          hipMemcpyDeviceToHost, S->stream));
    MI_HIP(hipStreamSynchronize(S->stream));
    return MI_OK;
}

int ensure_workspace(mi_solver* S)
{
    const size_t n = size_t(S->M->n());
    const bool mixed = S->cfg.precision == MI_PRECISION_MIXED;
    // Krylov vectors: CG uses slots 0-3 (r z p Ap), BiCGStab 0-7
    // (r rh p v s t y z); the FP32 pool adds 8 = outer residual, 9 = correction.
    // GMRES keeps its basis in its own workspace.
    const bool gmres = S->cfg.kind == MI_SOLVER_GMRES;
    const int nv = gmres ? 0 : S->cfg.kind == MI_SOLVER_BICGSTAB ? 8 : 4;
    for (int k = 0; k < nv; ++k)
        MI_TRY(S->w64[k].alloc(n));
    if (mixed)
        for (int k = 0; k < nv; ++k)
            MI_TRY(S->w32[k].alloc(n));
    if (mixed && !gmres) {
        MI_TRY(S->w32[8].alloc(n));
        MI_TRY(S->w32[9].alloc(n));
    }
    MI_TRY(S->st.alloc(1));
    MI_TRY(S->partials.alloc(size_t(std::max({kMaxReduceBlocks, kMaxStreamBlocks, kMaxSellBlocks})) *
                             kMaxReduceValues));
    MI_TRY(S->dscal.alloc(8));
    if (!S->h_st)
        MI_HIP(hipHostMalloc(reinterpret_cast<void**>(&S->h_st), sizeof(KState)));
    if (!S->h_scal)
        MI_HIP(hipHostMalloc(reinterpret_cast<void**>(&S->h_scal), 8 * sizeof(double)));
    const int hist = S->cfg.max_iterations > 0 ? S->cfg.max_iterations : 1;
    MI_TRY(S->hist.alloc(size_t(hist)));
    return MI_OK;
}

bool want_reorder(const mi_solver* S)
{
    const int r = S->cfg.reorder;
    return r > 0 || (r < 0 && S->A->n() >= 100000);
}

// Ordering of the working matrix: multicolour for the triangular-sweep
// preconditioners (ILU(0), SSOR) when asked (2) or by default at large sizes
// (their sweeps then have one level per colour instead of hundreds), reverse
// Cuthill-McKee otherwise.
bool want_colors(const mi_solver* S)
{
    const int r = S->cfg.reorder;
    const int pc = S->cfg.preconditioner;
    return r == 2 || (r < 0 && (pc == MI_PC_ILU || pc == MI_PC_SSOR));
}

// Computes the reverse Cuthill-McKee permutation of A's pattern and the
// reordered working matrix (once per pattern).
int build_reordering(mi_solver* S)
{
    mi_csr* A = S->A;
    const Pattern& src = *A->pat;
    auto ro = std::make_unique<Reordering>();
    ro->source = A->pat;
    const int n = int(src.n);
    if (want_colors(S))
        color_order(n, src.h_rp.data(), src.h_ci.data(), ro->perm, &ro->colors);
    else
        rcm_order(n, src.h_rp.data(), src.h_ci.data(), ro->perm);
    std::vector<int> brp, bci;
    permute_pattern(n, src.h_rp.data(), src.h_ci.data(), ro->perm, brp, bci, ro->vmap);
    ro->bandwidth_before = bandwidth(n, src.h_rp.data(), src.h_ci.data());
    ro->bandwidth_after = bandwidth(n, brp.data(), bci.data());
    MI_TRY(make_pattern(n, src.nnz, brp.data(), bci.data(), ro->pat, S->stream));
    MI_TRY(ro->d_perm.upload(ro->perm.data(), ro->perm.size(), S->stream));
    MI_TRY(ro->d_vmap.upload(ro->vmap.data(), ro->vmap.size(), S->stream));
    mi_csr* Mo = nullptr;
    MI_HIP(hipStreamSynchronize(S->stream));
    MI_TRY(make_csr_on_pattern(ro->pat, &Mo));
    S->Mown.reset(Mo);
    S->ro = std::move(ro);
    S->M = S->Mown.get();
    S->pc.reset();
    S->built_version = 0;
    return MI_OK;
}

int build(mi_solver* S)
{
    mi_csr* A = S->A;
    const bool mixed = S->cfg.precision == MI_PRECISION_MIXED;
    const auto t0 = Clock::now();
    if (want_reorder(S)) {
        if (!S->ro || S->ro->source != A->pat)
            MI_TRY(build_reordering(S));
    } else if (S->M != A) {
        S->ro.reset();
        S->Mown.reset();
        S->M = A;
        S->pc.reset();
        S->built_version = 0;
    }
    if (S->pc && S->built_version == A->version && !S->pending_amg)
        return MI_OK;
    mi_csr* M = S->M;
    if (S->ro) {
        // M = P A P^T: gather the values on the device (assembly leaves them there)
        const int64_t nnz = A->nnz();
        if (nnz) {
            k_gather_values<<<reduce_grid(nnz, 4), kBlock, 0, S->stream>>>(nnz, S->ro->d_vmap.p,
                                                                          A->val.p, M->val.p);
            MI_HIP(hipGetLastError());
        }
        M->values_changed(false);
        if (A->host_valid) {
            M->h_val.resize(size_t(nnz));
            for (int64_t k = 0; k < nnz; ++k)
                M->h_val[size_t(k)] = A->h_val[size_t(S->ro->vmap[size_t(k)])];
            M->host_valid = true;
        } else {
            MI_HIP(hipStreamSynchronize(S->stream));
            M->host_valid = false;
        }
    }
    if (mixed)
        MI_TRY(M->ensure_fp32_op(S->stream));
    if (S->pc && S->cfg.keep_preconditioner && !S->pending_amg) {
        // same pattern, new values: keep the preconditioner built on earlier values
        MI_HIP(hipStreamSynchronize(S->stream));
        S->built_version = A->version;
        S->setup_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
        return MI_OK;
    }
    if (!S->pc) {
        switch (S->cfg.preconditioner) {
        case MI_PC_AMG:
            S->pc = make_amg(S->cfg);
            break;
        case MI_PC_ILU:
            S->pc = make_ilu0();
            break;
        case MI_PC_SSOR:
            S->pc = make_ssor();
            break;
        default:
            S->pc = make_jacobi(); // Jacobi, None and unknown kinds (LinearSolvers.cpp:240-254)
        }
        S->pc->use_graph = S->cfg.use_graph != 0;
    }
    MI_HIP(hipStreamSynchronize(S->stream));
    if (S->pending_amg) {
        // an adopted hierarchy: it must have been built on this working pattern
        std::unique_ptr<AmgHierarchyHost> H = std::move(S->pending_amg);
        const HCsr& A0 = H->levels.front().A;
        if (A0.nr != int(M->n()) || A0.rp != M->pat->h_rp || A0.ci != M->pat->h_ci)
            return set_error(MI_ERR_BAD_ARG,
                             "adopted AMG hierarchy was built on another matrix pattern");
        MI_TRY(S->pc->install(*H, !mixed, mixed, S->stream));
    } else if (S->cfg.amg_resetup && S->pc->can_resetup(M, !mixed, mixed)) {
        // same pattern, new values: the hierarchy's values recomputed on the device
        MI_TRY(S->pc->resetup(M, !mixed, mixed, S->stream));
    } else {
        MI_TRY(S->pc->setup(M, !mixed, mixed, S->stream));
    }
    MI_TRY(ensure_workspace(S));
    MI_HIP(hipStreamSynchronize(S->stream));
    S->built_version = A->version;
    S->setup_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
    return MI_OK;
}

int check_interval(const mi_solver* S, int n)
{
    if (S->cfg.check_interval > 0)
        return S->cfg.check_interval;
    return n < 200000 ? 32 : 8;
}

template <class T> int precond(Precond* pc, const T* r, T* z, hipStream_t s)
{
    if constexpr (sizeof(T) == 8)
        return pc->apply64(r, z, s);
    else
        return pc->apply32(r, z, s);
}

template <class T>
int precond_cg(Precond* pc, const CgFuse<T>& f, T* z, unsigned* grr, unsigned* grz, hipStream_t s)
{
    if constexpr (sizeof(T) == 8)
        return pc->apply_cg64(f, z, grr, grz, s);
    else
        return pc->apply_cg32(f, z, grr, grz, s);
}

// Polls KState every `chunk` issued iterations until the device says done.
// Issues iterations in chunks between polls of the device state.  Kernels of
// an iteration after convergence return at once, but a preconditioner apply
// (a V-cycle, the ILU sweeps) runs regardless, so each chunk is sized from
// the convergence rate observed between the last two polls to the
// iterations still needed (at most `chunk`).
template <class Body> int iterate(mi_solver* S, KState* st, int max_iter, int chunk, Body&& body)
{
    int issued = 0;
    int prev_iter = -1;
    double prev_last = 0.0;
    int next = std::min(chunk, 2);
    for (;;) {
        MI_HIP(hipMemcpyAsync(S->h_st, st, sizeof(KState), hipMemcpyDeviceToHost, S->stream));
        MI_HIP(hipStreamSynchronize(S->stream));
        const KState& h = *S->h_st;
        if (h.done || issued >= max_iter)
            break;
        if (prev_iter >= 0 && h.iter > prev_iter && h.last > 0.0 && prev_last > 0.0) {
            const double rho = std::pow(h.last / prev_last, 1.0 / double(h.iter - prev_iter));
            next = chunk;
            if (rho > 0.0 && rho < 1.0 && h.tol > 0.0 && h.tol < h.last) {
                const double need = std::log(h.tol / h.last) / std::log(rho);
                if (need < double(chunk))
                    next = std::max(1, int(std::ceil(need)));
            }
        }
        prev_iter = h.iter;
        prev_last = h.last;
        const int k = std::min(next, max_iter - issued);
        for (int i = 0; i < k; ++i)
            MI_TRY(body());
        issued += k;
    }
    return MI_OK;
}

// One Krylov solve A x = b in element type T on the working matrix with the
// vector pool w; the final KState is left in S->h_st.  With xacc set (FP32
// inner solves) the iterate is added to that FP64 vector instead of x.
// The working matrix as the Krylov kernels stream it: FP64 values, or the
// FP32 inner-iteration operator.
static int krylov_view(mi_solver* S, StreamMat<double>& V)
{
    return stream_view64(S->M, S->stream, V, true);
}
static int krylov_view(mi_solver* S, StreamMat<float>& V)
{
    return stream_view32(S->M, S->stream, true, V, true);
}

template <class T>
int run_krylov(mi_solver* S, const T* invD, const T* b, T* x, DevBuf<T>* w, double tol,
               int max_iter, bool warm, double* xacc = nullptr)
{
    const Pattern& P = *S->M->pat;
    hipStream_t s = S->stream;
    Precond* pc = S->pc.get();
    const int n = int(P.n);
    StreamMat<T> V;
    MI_TRY(krylov_view(S, V));
    const T* val = V.val;
    const int chunk = check_interval(S, n);
    if (S->cfg.kind == MI_SOLVER_BICGSTAB) {
        BiArgs<T> a{};
        a.n = n;
        a.nnz = P.nnz;
        a.rp = P.rp.p;
        a.ci = P.ci.p;
        a.val = val;
        a.b = b;
        a.x = x;
        a.r = w[0].p;
        a.rh = w[1].p;
        a.p = w[2].p;
        a.v = w[3].p;
        a.s = w[4].p;
        a.t = w[5].p;
        a.y = w[6].p;
        a.z = w[7].p;
        a.st = S->st.p;
        a.partials = S->partials.p;
        a.history = S->hist.p;
        a.lanes = P.lanes;
        a.grid = range_grid(n, 256);
        a.blk = P.rb.blk.p;
        a.diag = V.diag;
        a.cw = V.cw;
        a.hdr = V.hdr;
        a.soff = V.soff;
        a.eci = V.eci;
        a.ecd = V.ecd;
        a.nsl = V.nsl;
        a.nrb = P.rb.nrb;
        a.xacc = xacc;
        MI_TRY(bicg_init<T>(a, warm, tol, max_iter, s));
        return iterate(S, a.st, max_iter, chunk, [&]() -> int {
            MI_TRY(bicg_stage1<T>(a, s));
            MI_TRY(precond<T>(pc, a.p, a.y, s));
            MI_TRY(bicg_stage2<T>(a, s));
            MI_TRY(precond<T>(pc, a.s, a.z, s));
            return bicg_stage3<T>(a, s);
        });
    }
    CgArgs<T> a{};
    a.n = n;
    a.nnz = P.nnz;
    a.rp = P.rp.p;
    a.ci = P.ci.p;
    a.val = val;
    a.invD = invD;
    a.b = b;
    a.x = x;
    a.r = w[0].p;
    a.z = w[1].p;
    a.p = w[2].p;
    a.Ap = w[3].p;
    a.st = S->st.p;
    a.partials = S->partials.p;
    a.history = S->hist.p;
    a.lanes = P.lanes;
    a.grid = range_grid(n, 256);
    a.blk = P.rb.blk.p;
    a.nrb = P.rb.nrb;
    a.diag = V.diag;
    a.cw = V.cw;
    a.hdr = V.hdr;
    a.soff = V.soff;
    a.eci = V.eci;
    a.ecd = V.ecd;
    a.nsl = V.nsl;
    a.xacc = xacc;
    const bool jacobi = invD != nullptr;
    if (jacobi) {
        MI_TRY(cg_init_jacobi<T>(a, warm, tol, max_iter, s));
    } else {
        MI_TRY(cg_init_r<T>(a, warm, s));
        MI_TRY(precond<T>(pc, a.r, a.z, s));
        MI_TRY(cg_init_p<T>(a, tol, max_iter, s));
    }
    // With an AMG V-cycle the CG update and the r.z reduction run inside the
    // cycle's first and last launches (two launches and two vector passes
    // fewer per iteration); the r.r and r.z partials go to their own
    // sections of the partials buffer.
    const bool fused = !jacobi && pc->cg_fusable(sizeof(T) == 8);
    constexpr size_t kSection = size_t(std::max({kMaxReduceBlocks, kMaxStreamBlocks, kMaxSellBlocks}));
    static_assert(3 * kSection <= kSection * kMaxReduceValues, "partials sections");
    const CgFuse<T> f{a.r, a.p, a.Ap, a.xacc ? nullptr : a.x, a.xacc, a.st,
                      a.partials + kSection, a.partials + 2 * kSection};
    return iterate(S, a.st, max_iter, chunk, [&]() -> int {
        if (jacobi)
            return cg_iteration_jacobi<T>(a, s);
        if (fused) {
            unsigned grr = 0, grz = 0;
            MI_TRY(cg_iteration_spmv<T>(a, s));
            MI_TRY(precond_cg<T>(pc, f, a.z, &grr, &grz, s));
            return cg_iteration_fused_tail<T>(a, f.prr, grr, f.prz, grz, s);
        }
        MI_TRY(cg_iteration_part1<T>(a, s));
        MI_TRY(precond<T>(pc, a.r, a.z, s));
        return cg_iteration_part2<T>(a, s);
    });
}

void copy_history(mi_solver* S, int count)
{
    if (count <= 0)
        return;
    const size_t old = S->history.size();
    S->history.resize(old + size_t(count));
    (void)hipMemcpy(S->history.data() + old, S->hist.p, sizeof(double) * size_t(count),
                    hipMemcpyDeviceToHost);
}

int solve_fp64(mi_solver* S, const double* b, double* x, mi_stats* out)
{
    const mi_solver_cfg& c = S->cfg;
    Precond* pc = S->pc.get();
    const double* invD = c.kind == MI_SOLVER_CG ? pc->inv_diag64() : nullptr;
    const int n = int(S->M->n());
    bool warm = c.warm_start != 0;
    int total = 0, outer = 0;
    int status = MI_OK;
    double rel_true = 0.0;
    MI_TRY(norm2(S, n, b, S->dscal.p + 1));
    if (warm)
        MI_TRY(true_residual(S, b, x, S->w64[0].p, nullptr, S->dscal.p + 2));
    MI_TRY(read_scalars(S, 3));
    const double bnorm = std::max(S->h_scal[1], 1e-30);
    const double init_res = warm ? S->h_scal[2] : S->h_scal[1];
    for (;;) {
        MI_TRY(run_krylov<double>(S, invD, b, x, S->w64, c.tolerance, c.max_iterations - total,
                                  warm));
        const KState st = *S->h_st;
        copy_history(S, st.iter);
        total += st.iter;
        ++outer;
        S->last = st.last;
        MI_TRY(true_residual(S, b, x, nullptr, nullptr, S->dscal.p));
        MI_TRY(read_scalars(S, 1));
        rel_true = S->h_scal[0] / bnorm;
        if (st.done == 3) {
            status = (st.last < c.tolerance) ? MI_OK : MI_ERR_BREAKDOWN;
            break;
        }
        if (st.done != 1) {
            status = MI_ERR_NOT_CONVERGED;
            break;
        }
        if (!c.true_residual || rel_true < c.tolerance)
            break;
        if (total >= c.max_iterations || outer >= std::max(c.max_outer, 1)) {
            status = MI_ERR_NOT_CONVERGED;
            break;
        }
        // The recurrence residual met the tolerance but the true residual did
        // not: restart from the current x (FP64 refinement).
        MI_TRY(true_residual(S, b, x, S->w64[0].p, nullptr, S->dscal.p));
        warm = true;
    }
    if (c.true_residual)
        S->last = rel_true;
    out->iterations = total;
    out->outer_iterations = outer;
    out->status = status;
    out->converged = status == MI_OK;
    out->initial_residual = init_res;
    out->relative_residual = S->last;
    out->final_residual = S->last * bnorm;
    out->true_relative_residual = rel_true;
    return status;
}

int solve_mixed(mi_solver* S, const double* b, double* x, mi_stats* out)
{
    const mi_solver_cfg& c = S->cfg;
    Precond* pc = S->pc.get();
    const float* invD = c.kind == MI_SOLVER_CG ? pc->inv_diag32() : nullptr;
    const int n = int(S->M->n());
    hipStream_t s = S->stream;
    float* r32 = S->w32[8].p; // outer residual b - A x in FP32
    float* d32 = S->w32[9].p; // unused FP32 iterate slot of the inner solve
    const bool warm = c.warm_start != 0;
    if (!warm)
        MI_HIP(hipMemsetAsync(x, 0, sizeof(double) * size_t(n), s));
    MI_TRY(norm2(S, n, b, S->dscal.p + 1));
    MI_TRY(true_residual(S, b, x, nullptr, r32, S->dscal.p));
    MI_TRY(read_scalars(S, 2));
    const double bnorm = std::max(S->h_scal[1], 1e-30);
    const double init_res = S->h_scal[0];
    double rn = init_res;
    int total = 0, outer = 0, status = MI_OK;
    const double inner_tol = c.inner_tolerance > 0 ? c.inner_tolerance : 1e-4;
    // `!(x < tol)` so that a NaN residual never counts as converged, as the
    // reference's `last_ < tol` test (src/solver/LinearSolvers.cpp:323)
    while (!(rn / bnorm < c.tolerance)) {
        if (!std::isfinite(rn)) {
            // a NaN/Inf in b or an overflowing inner solve: nothing to refine
            status = set_error(MI_ERR_BREAKDOWN, "mixed refinement: non-finite residual %g",
                               rn);
            break;
        }
        if (outer >= std::max(c.max_outer, 1) || total >= c.max_iterations) {
            status = MI_ERR_NOT_CONVERGED;
            break;
        }
        // inner: A d = r32 from d = 0 to inner_tol relative to ||r32||, or
        // less when that already meets the outer target; x += d is done by
        // the inner update kernels in FP64 (xacc)
        const double tol_in = std::max(inner_tol, 0.5 * c.tolerance * bnorm / rn);
        MI_TRY(run_krylov<float>(S, invD, r32, d32, S->w32, tol_in, c.max_iterations - total,
                                 false, x));
        const KState st = *S->h_st;
        copy_history(S, st.iter);
        total += st.iter;
        ++outer;
        MI_TRY(true_residual(S, b, x, nullptr, r32, S->dscal.p));
        MI_TRY(read_scalars(S, 1));
        rn = S->h_scal[0];
        if (st.done == 3 && st.iter == 0) {
            status = MI_ERR_BREAKDOWN;
            break;
        }
    }
    S->last = rn / bnorm;
    out->iterations = total;
    out->outer_iterations = outer;
    out->status = status;
    out->converged = status == MI_OK;
    out->initial_residual = init_res;
    out->final_residual = rn;
    out->relative_residual = S->last;
    out->true_relative_residual = S->last;
    return status;
}

// ------------------------------------------------------------- GMRES(m)
int ensure_gmres(mi_solver* S)
{
    const int m = std::max(1, S->cfg.restart);
    if (m > kGmMaxRestart)
        return set_error(MI_ERR_BAD_ARG, "GMRES restart %d exceeds the device limit %d", m,
                         kGmMaxRestart);
    if (!S->gm)
        S->gm = std::make_unique<GmWork>();
    GmWork& W = *S->gm;
    const int64_t n = S->M->n();
    const bool mixed = S->cfg.precision == MI_PRECISION_MIXED;
    if (W.m == m && W.n == n && (mixed ? W.V32.p != nullptr : W.V64.p != nullptr))
        return MI_OK;
    W.m = m;
    W.n = n;
    W.ld = size_t((n + 63) / 64 * 64); // 512-byte aligned basis vectors
    W.G = gm_grid(n);
    const size_t nb = size_t(m + 1) * W.ld;
    if (mixed) {
        W.V64.release();
        W.w64.release();
        MI_TRY(W.V32.alloc(nb));
        MI_TRY(W.w32.alloc(size_t(n)));
    } else {
        W.V32.release();
        W.w32.release();
        MI_TRY(W.V64.alloc(nb));
        MI_TRY(W.w64.alloc(size_t(n)));
    }
    MI_TRY(W.small.alloc(gm_small_size(m)));
    MI_TRY(W.partials.alloc(size_t(m + 1) * W.G));
    MI_TRY(W.st.alloc(1));
    if (!W.h_st)
        MI_HIP(hipHostMalloc(reinterpret_cast<void**>(&W.h_st), sizeof(GmState)));
    return MI_OK;
}

// Restarted GMRES(m) of GMRESmSolver::solve (src/solver/LinearSolvers.cpp:266-340)
// on the working matrix, from the iterate in x (FP64).  T is the precision
// of the Arnoldi steps; every cycle starts from the FP64 residual b - A x.
// With `mixed` the in-cycle estimate only ends a cycle and the FP64 residual
// test at the next cycle start decides convergence.  The final state is left
// in the pinned mirror of the workspace.
template <class T>
int run_gmres(mi_solver* S, const double* b, double* x, bool mixed, int* cycles)
{
    GmWork& W = *S->gm;
    mi_csr* M = S->M;
    const Pattern& P = *M->pat;
    hipStream_t s = S->stream;
    Precond* pc = S->pc.get();
    const int n = int(P.n), m = W.m;
    GmArgs<T> a{};
    a.n = n;
    a.m = m;
    a.ld = W.ld;
    MI_TRY(stream_view64(M, s, a.A64, true));
    if constexpr (sizeof(T) == 8) {
        a.V = W.V64.p;
        a.w = W.w64.p;
        a.invD = pc->inv_diag64();
        a.A = a.A64;
    } else {
        a.V = W.V32.p;
        a.w = W.w32.p;
        a.invD = pc->inv_diag32();
        MI_TRY(stream_view32(M, s, true, a.A, true));
    }
    a.b = b;
    a.x = x;
    a.g = W.st.p;
    double* sm = W.small.p;
    a.H = sm;
    sm += size_t(m + 1) * size_t(m);
    a.cs = sm;
    sm += m;
    a.sn = sm;
    sm += m;
    a.y = sm;
    sm += m;
    a.sv = sm;
    sm += m + 1;
    a.h1 = sm;
    sm += m + 1;
    a.h2 = sm;
    a.partials = W.partials.p;
    a.history = S->hist.p;
    a.hist_cap = int(S->hist.n);
    a.G = int(W.G);
    a.Gs = launch_grid(a.A64); // A64 and A share the pattern's form
    MI_TRY(gm_init<T>(a, S->dscal.p + 1, S->cfg.tolerance, S->cfg.max_iterations, mixed, s));
    const int chunk = check_interval(S, n);
    auto poll = [&]() -> int {
        MI_HIP(hipMemcpyAsync(W.h_st, W.st.p, sizeof(GmState), hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
        return MI_OK;
    };
    for (;;) {
        // cycle start: v_0 = M^-1 (b - A x), beta, tests, v_0 /= beta
        MI_TRY(gm_cycle_residual<T>(a, s));
        if (!a.invD)
            MI_TRY(precond<T>(pc, a.w, a.V, s));
        MI_TRY(gm_cycle_begin<T>(a, s));
        MI_TRY(poll());
        if (W.h_st->k.done)
            break;
        ++*cycles;
        const int room = std::min(m, W.h_st->k.max_iter - W.h_st->k.iter);
        // Arnoldi steps in groups between polls, sized from the observed
        // convergence rate as in the CG / BiCGStab driver
        int j = 0, prev_j = -1, next = std::min(chunk, 2);
        double prev_last = 0.0;
        while (j < room) {
            const int k = std::min(next, room - j);
            for (int q = 0; q < k; ++q, ++j) {
                MI_TRY(gm_step_spmv<T>(a, j, s));
                if (!a.invD)
                    MI_TRY(precond<T>(pc, a.w, a.V + size_t(j + 1) * a.ld, s));
                MI_TRY(gm_step_orth<T>(a, j, s));
            }
            MI_TRY(poll());
            const GmState& h = *W.h_st;
            if (h.k.done || h.cycle_end)
                break;
            next = chunk;
            if (prev_j >= 0 && h.j > prev_j && h.k.last > 0.0 && prev_last > 0.0) {
                const double rho = std::pow(h.k.last / prev_last, 1.0 / double(h.j - prev_j));
                if (rho > 0.0 && rho < 1.0 && h.k.tol > 0.0 && h.k.tol < h.k.last) {
                    const double need = std::log(h.k.tol / h.k.last) / std::log(rho);
                    if (need < double(chunk))
                        next = std::max(1, int(std::ceil(need)));
                }
            }
            prev_j = h.j;
            prev_last = h.k.last;
        }
        // cycle end: y = H^-1 s, x += V y
        MI_TRY(gm_cycle_end<T>(a, s));
        if (W.h_st->k.done) {
            MI_HIP(hipStreamSynchronize(s));
            break;
        }
    }
    return MI_OK;
}

int solve_gmres(mi_solver* S, const double* b, double* x, mi_stats* out)
{
    const mi_solver_cfg& c = S->cfg;
    const bool mixed = c.precision == MI_PRECISION_MIXED;
    const int n = int(S->M->n());
    const bool warm = c.warm_start != 0;
    MI_TRY(ensure_gmres(S));
    if (!warm)
        MI_HIP(hipMemsetAsync(x, 0, sizeof(double) * size_t(n), S->stream));
    MI_TRY(norm2(S, n, b, S->dscal.p + 1));
    if (warm)
        MI_TRY(true_residual(S, b, x, nullptr, nullptr, S->dscal.p + 2));
    MI_TRY(read_scalars(S, 3));
    const double bnorm = std::max(S->h_scal[1], 1e-30);
    const double init_res = warm ? S->h_scal[2] : S->h_scal[1];
    int cycles = 0;
    if (mixed)
        MI_TRY(run_gmres<float>(S, b, x, true, &cycles));
    else
        MI_TRY(run_gmres<double>(S, b, x, false, &cycles));
    const GmState h = *S->gm->h_st;
    copy_history(S, std::min<int>(h.steps, int(S->hist.n)));
    MI_TRY(true_residual(S, b, x, nullptr, nullptr, S->dscal.p));
    MI_TRY(read_scalars(S, 1));
    const int status = h.k.done == 1   ? MI_OK
                       : h.k.done == 3 ? MI_ERR_BREAKDOWN
                                       : MI_ERR_NOT_CONVERGED;
    S->last = h.k.last;
    out->iterations = h.k.iter;
    out->outer_iterations = cycles;
    out->status = status;
    out->converged = status == MI_OK;
    out->initial_residual = init_res;
    out->relative_residual = h.k.last;
    out->final_residual = h.k.last * bnorm;
    out->true_relative_residual = S->h_scal[0] / bnorm;
    return status;
}

int solve_device(mi_solver* S, const double* b, double* x, mi_stats* stats)
{
    mi_stats local{};
    mi_stats* out = stats ? stats : &local;
    std::memset(out, 0, sizeof(*out));
    if (!S->A)
        return set_error(MI_ERR_BAD_ARG, "solver has no matrix: call mi_solver_setup first");
    if (S->cfg.kind != MI_SOLVER_CG && S->cfg.kind != MI_SOLVER_BICGSTAB &&
        S->cfg.kind != MI_SOLVER_GMRES)
        return set_error(MI_ERR_UNSUPPORTED, "solver kind %d not available on device",
                         S->cfg.kind);
    MI_TRY(build(S));
    out->setup_seconds = S->setup_seconds;
    S->history.clear();
    if (S->A->n() == 0) {
        out->converged = 1;
        S->last = 0.0;
        return MI_OK;
    }
    const auto t0 = Clock::now();
    const int n = int(S->A->n());
    const double* bw = b;
    double* xw = x;
    if (S->ro) {
        // solve in the reordered numbering: b_w = P b, x_w = P x
        MI_TRY(S->bw.alloc(size_t(n)));
        MI_TRY(S->xw.alloc(size_t(n)));
        k_gather<<<grid_for(n, kBlock), kBlock, 0, S->stream>>>(n, S->ro->d_perm.p, b, S->bw.p);
        if (S->cfg.warm_start)
            k_gather<<<grid_for(n, kBlock), kBlock, 0, S->stream>>>(n, S->ro->d_perm.p, x, S->xw.p);
        MI_HIP(hipGetLastError());
        bw = S->bw.p;
        xw = S->xw.p;
    }
    int rc;
    if (S->cfg.kind == MI_SOLVER_GMRES)
        rc = solve_gmres(S, bw, xw, out);
    else if (S->cfg.precision == MI_PRECISION_MIXED)
        rc = solve_mixed(S, bw, xw, out);
    else
        rc = solve_fp64(S, bw, xw, out);
    if (rc != MI_OK && rc != MI_ERR_NOT_CONVERGED && rc != MI_ERR_BREAKDOWN)
        return rc;
    if (S->ro) {
        k_scatter<<<grid_for(n, kBlock), kBlock, 0, S->stream>>>(n, S->ro->d_perm.p, S->xw.p, x);
        MI_HIP(hipGetLastError());
    }
    MI_HIP(hipStreamSynchronize(S->stream));
    out->solve_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
    return rc;
}

} // namespace

namespace mi
{

int solver_adopt_amg(mi_solver* S, const AmgHierarchyHost& H)
{
    if (!S || !S->A)
        return set_error(MI_ERR_BAD_ARG, "solver has no matrix: call mi_solver_setup first");
    if (S->cfg.preconditioner != MI_PC_AMG || !S->cfg.keep_preconditioner)
        return set_error(MI_ERR_UNSUPPORTED,
                         "only an AMG solver that keeps its preconditioner adopts a hierarchy");
    if (H.levels.empty())
        return set_error(MI_ERR_BAD_ARG, "empty hierarchy");
    S->pending_amg = std::make_unique<AmgHierarchyHost>(H);
    return MI_OK;
}

const AmgHierarchyHost* solver_amg(const mi_solver* S)
{
    return S && S->pc ? S->pc->hierarchy() : nullptr;
}

} // namespace mi

extern "C" {

MI_API void mi_solver_cfg_default(mi_solver_cfg* c)
{
    if (!c)
        return;
    std::memset(c, 0, sizeof(*c));
    c->kind = MI_SOLVER_CG;
    c->preconditioner = MI_PC_JACOBI;
    c->tolerance = 1.0e-8;
    c->max_iterations = 1000;
    c->restart = 30;
    c->precision = MI_PRECISION_FP64;
    c->warm_start = 0;
    c->true_residual = 0;
    c->inner_tolerance = 1.0e-4;
    c->max_outer = 20;
    c->reorder = -1;
    c->check_interval = 0;
    c->amg_max_levels = 10;
    c->amg_coarse_size = 2048;
    c->amg_strength = 0.02;
    c->amg_jacobi_omega = 4.0 / 3.0;
    c->amg_truncation = 0.1;
    c->amg_smoother = 0;
    c->amg_sweeps = 2;
    c->amg_coarse_sweeps = 1;
    c->keep_preconditioner = 0;
    c->use_graph = 1;
    c->amg_resetup = 1;
}

MI_API int mi_solver_create(const mi_solver_cfg* cfg, mi_solver** out)
{
    clear_error();
    if (!out)
        return set_error(MI_ERR_BAD_ARG, "null output handle");
    *out = nullptr;
    MI_TRY(ensure_device());
    auto* S = new (std::nothrow) mi_solver();
    if (!S)
        return set_error(MI_ERR_DEVICE, "out of host memory");
    if (cfg)
        S->cfg = *cfg;
    else
        mi_solver_cfg_default(&S->cfg);
    if (S->cfg.max_iterations < 0 || !(S->cfg.tolerance >= 0.0)) {
        delete S;
        return set_error(MI_ERR_BAD_ARG, "bad tolerance or iteration limit");
    }
    if (S->cfg.kind == MI_SOLVER_GMRES && S->cfg.restart > kGmMaxRestart) {
        const int m = S->cfg.restart;
        delete S;
        return set_error(MI_ERR_BAD_ARG, "GMRES restart %d exceeds the device limit %d", m,
                         kGmMaxRestart);
    }
    if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) {
        delete S;
        return set_error(MI_ERR_DEVICE, "hipStreamCreate failed");
    }
    *out = S;
    return MI_OK;
}

MI_API int mi_solver_setup(mi_solver* S, mi_csr* A)
{
    clear_error();
    if (!S || !A)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    if (S->A != A) {
        S->A = A;
        S->pc.reset();
        S->built_version = 0;
    }
    return build(S);
}

MI_API int mi_solver_solve_device(mi_solver* S, const double* b, double* x, mi_stats* stats)
{
    clear_error();
    if (!S || (S->A && S->A->n() && (!b || !x)))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    return solve_device(S, b, x, stats);
}

MI_API int mi_solver_solve(mi_solver* S, const double* b, double* x, mi_stats* stats)
{
    clear_error();
    if (!S || !S->A)
        return set_error(MI_ERR_BAD_ARG, "solver has no matrix: call mi_solver_setup first");
    const size_t n = size_t(S->A->n());
    if (n && (!b || !x))
        return set_error(MI_ERR_BAD_ARG, "null vector");
    MI_TRY(ensure_device());
    MI_TRY(S->b.alloc(n));
    MI_TRY(S->x.alloc(n));
    if (n) {
        MI_HIP(hipMemcpyAsync(S->b.p, b, n * sizeof(double), hipMemcpyHostToDevice, S->stream));
        if (S->cfg.warm_start)
            MI_HIP(hipMemcpyAsync(S->x.p, x, n * sizeof(double), hipMemcpyHostToDevice, S->stream));
    }
    const int rc = solve_device(S, S->b.p, S->x.p, stats);
    if (rc != MI_OK && rc != MI_ERR_NOT_CONVERGED && rc != MI_ERR_BREAKDOWN)
        return rc;
    if (n) {
        MI_HIP(hipMemcpyAsync(x, S->x.p, n * sizeof(double), hipMemcpyDeviceToHost, S->stream));
        MI_HIP(hipStreamSynchronize(S->stream));
    }
    return rc;
}

MI_API int mi_solver_precond_apply(mi_solver* S, const double* r, double* z)
{
    clear_error();
    if (!S || !S->A)
        return set_error(MI_ERR_BAD_ARG, "solver has no matrix: call mi_solver_setup first");
    const int n = int(S->A->n());
    if (n && (!r || !z))
        return set_error(MI_ERR_BAD_ARG, "null vector");
    MI_TRY(ensure_device());
    MI_TRY(build(S));
    if (n == 0)
        return MI_OK;
    hipStream_t s = S->stream;
    MI_TRY(S->b.alloc(size_t(n)));
    MI_TRY(S->x.alloc(size_t(n)));
    MI_TRY(S->bw.alloc(size_t(n)));
    MI_TRY(S->xw.alloc(size_t(n)));
    MI_HIP(hipMemcpyAsync(S->b.p, r, size_t(n) * sizeof(double), hipMemcpyHostToDevice, s));
    const double* rw = S->b.p;
    double* zw = S->ro ? S->xw.p : S->x.p;
    if (S->ro) {
        k_gather<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, S->ro->d_perm.p, S->b.p, S->bw.p);
        rw = S->bw.p;
    }
    if (S->cfg.precision == MI_PRECISION_MIXED) {
        DevBuf<float> r32, z32;
        MI_TRY(r32.alloc(size_t(n)));
        MI_TRY(z32.alloc(size_t(n)));
        k_to_f32<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, rw, r32.p);
        MI_TRY(S->pc->apply32(r32.p, z32.p, s));
        k_to_f64<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, z32.p, zw);
        MI_HIP(hipStreamSynchronize(s));
    } else {
        MI_TRY(S->pc->apply64(rw, zw, s));
    }
    if (S->ro)
        k_scatter<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, S->ro->d_perm.p, S->xw.p, S->x.p);
    MI_HIP(hipGetLastError());
    MI_HIP(hipMemcpyAsync(z, S->x.p, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

MI_API int mi_solver_info(const mi_solver* S, int* levels, double* operator_complexity,
                          int* bandwidth_before, int* bandwidth_after)
{
    if (!S)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    if (levels)
        *levels = S->pc ? S->pc->levels() : 0;
    if (operator_complexity)
        *operator_complexity = S->pc ? S->pc->complexity() : 0.0;
    if (bandwidth_before)
        *bandwidth_before = S->ro ? S->ro->bandwidth_before : -1;
    if (bandwidth_after)
        *bandwidth_after = S->ro ? S->ro->bandwidth_after : -1;
    return MI_OK;
}

MI_API int mi_solver_amg_host(mi_solver* S, mi_amg_host** out)
{
    clear_error();
    if (!S || !out)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    *out = nullptr;
    if (!S->pc)
        return set_error(MI_ERR_UNSUPPORTED, "the solver has built no preconditioner");
    MI_TRY(ensure_device());
    AmgHierarchyHost H;
    MI_TRY(S->pc->host_copy(H, S->stream));
    return amg_host_wrap(std::move(H), out);
}

MI_API int mi_solver_time_spmv(mi_solver* S, int precision, int reps, double* avg_us,
                               double* bytes)
{
    clear_error();
    if (!S || !S->A || reps < 1 || !avg_us)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    MI_TRY(ensure_device());
    MI_TRY(build(S));
    mi_csr* M = S->M;
    const int64_t n = M->n(), nnz = M->nnz();
    const size_t vb = precision == MI_PRECISION_MIXED ? 4 : 8;
    if (bytes)
        *bytes = double(nnz) * double(vb + 4) + double(n + 1) * 4.0 + 2.0 * double(n) * double(vb);
    if (n == 0) {
        *avg_us = 0.0;
        return MI_OK;
    }
    DevBuf<char> xb, yb;
    MI_TRY(xb.alloc(size_t(n) * vb));
    MI_TRY(yb.alloc(size_t(n) * vb));
    MI_HIP(hipMemsetAsync(xb.p, 0, size_t(n) * vb, default_stream()));
    MI_HIP(hipStreamSynchronize(default_stream()));
    return mi_csr_time_spmv(M, precision, xb.p, yb.p, reps, avg_us);
}

MI_API int mi_solver_time_spmv_cold(mi_solver* S, int precision, int reps, int64_t flush_bytes,
                                    double* avg_us, double* bytes)
{
    clear_error();
    if (!S || !S->A || reps < 1 || !avg_us)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    MI_TRY(ensure_device());
    MI_TRY(build(S));
    mi_csr* M = S->M;
    const int64_t n = M->n(), nnz = M->nnz();
    const size_t vb = precision == MI_PRECISION_MIXED ? 4 : 8;
    if (bytes)
        *bytes = double(nnz) * double(vb + 4) + double(n + 1) * 4.0 + 2.0 * double(n) * double(vb);
    if (n == 0) {
        *avg_us = 0.0;
        return MI_OK;
    }
    DevBuf<char> xb, yb;
    MI_TRY(xb.alloc(size_t(n) * vb));
    MI_TRY(yb.alloc(size_t(n) * vb));
    MI_HIP(hipMemsetAsync(xb.p, 0, size_t(n) * vb, default_stream()));
    MI_HIP(hipStreamSynchronize(default_stream()));
    return mi_csr_time_spmv_cold(M, precision, xb.p, yb.p, reps, flush_bytes, avg_us);
}

MI_API int mi_adjoint_solve(mi_solver* S, const double* g, double* lambda, mi_stats* stats)
{
    clear_error();
    if (!S || !S->A)
        return set_error(MI_ERR_BAD_ARG, "solver has no matrix: call mi_solver_setup first");
    if (S->A->n() && (!g || !lambda))
        return set_error(MI_ERR_BAD_ARG, "null vector");
    MI_TRY(ensure_device());
    if (!S->At || !S->adj || S->at_src != S->A) {
        S->adj.reset();
        S->At.reset();
        mi_csr* T = nullptr;
        MI_TRY(mi_csr_transpose(S->A, &T));
        S->At.reset(T);
        mi_solver* a = nullptr;
        MI_TRY(mi_solver_create(&S->cfg, &a));
        S->adj.reset(a);
        MI_TRY(mi_solver_setup(a, T));
        S->at_src = S->A;
        S->at_version = S->A->version;
    } else if (S->at_version != S->A->version) {
        MI_TRY(mi_csr_transpose_values(S->A, S->At.get()));
        S->at_version = S->A->version;
    }
    return solve_device(S->adj.get(), g, lambda, stats);
}

MI_API double mi_solver_last_residual(const mi_solver* S) { return S ? S->last : 0.0; }

MI_API int mi_solver_history(const mi_solver* S, double* out, int capacity, int* count)
{
    if (!S)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    const int n = int(S->history.size());
    if (count)
        *count = n;
    if (out)
        std::memcpy(out, S->history.data(), sizeof(double) * size_t(std::min(n, capacity)));
    return MI_OK;
}

MI_API int mi_solver_destroy(mi_solver* S)
{
    if (S && S->stream)
        (void)hipStreamSynchronize(S->stream);
    delete S;
    return MI_OK;
}

} // extern "C"
