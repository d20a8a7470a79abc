// amg_host_api.cpp — host-only entry points: the SA-AMG hierarchy and the
// RCM permutation the solver builds during setup, exposed for inspection
// (no device is touched).
#include "amg_setup.hpp"
#include "checkpoint.hpp"
#include "mi_common.hpp"
#include "reorder.hpp"

#include <cstring>
#include <memory>

struct mi_amg_host
{
    mi::AmgHierarchyHost h;
};

using namespace mi;

namespace mi
{
const AmgHierarchyHost* amg_host_hierarchy(const mi_amg_host* h)
{
    return h ? &h->h : nullptr;
}

int amg_host_wrap(AmgHierarchyHost&& H, mi_amg_host** out)
{
    auto w = std::make_unique<mi_amg_host>();
    w->h = std::move(H);
    *out = w.release();
    return MI_OK;
}
} // namespace mi

extern "C" {

// rowPtr starts at 0 and never decreases; every column index in [0, n).
static int check_host_csr(int64_t n, const int32_t* rp, const int32_t* ci)
{
    if (n == 0)
        return MI_OK;
    if (rp[0] != 0)
        return set_error(MI_ERR_BAD_ARG, "rowPtr[0] = %d, expected 0", rp[0]);
    for (int64_t i = 0; i < n; ++i) {
        if (rp[i + 1] < rp[i])
            return set_error(MI_ERR_BAD_ARG, "rowPtr decreases at row %lld", (long long)i);
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] < 0 || ci[k] >= n)
                return set_error(MI_ERR_BAD_ARG, "column %d out of range in row %lld", ci[k],
                                 (long long)i);
    }
    return MI_OK;
}

MI_API int mi_amg_host_build(int64_t n, const int32_t* rp, const int32_t* ci, const double* v,
                             const mi_solver_cfg* cfg, mi_amg_host** out)
{
    clear_error();
    if (!out || n < 0 || n > INT32_MAX - 1 || (n > 0 && (!rp || !ci || !v)))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    *out = nullptr;
    MI_TRY(check_host_csr(n, rp, ci));
    HCsr A;
    A.nr = A.nc = int(n);
    A.rp.assign(rp, rp + n + 1);
    const int64_t nnz = n > 0 ? rp[n] : 0;
    A.ci.assign(ci, ci + nnz);
    A.v.assign(v, v + nnz);
    for (int64_t k = 0; k < nnz; ++k)
        if (A.ci[size_t(k)] < 0 || A.ci[size_t(k)] >= n)
            return set_error(MI_ERR_BAD_ARG, "column out of range");
    AmgParams p;
    if (cfg) {
        if (cfg->amg_max_levels > 0)
            p.max_levels = cfg->amg_max_levels;
        if (cfg->amg_coarse_size > 0)
            p.coarse_size = cfg->amg_coarse_size;
        if (cfg->amg_strength >= 0)
            p.theta = cfg->amg_strength;
        if (cfg->amg_jacobi_omega > 0)
            p.omega_scale = cfg->amg_jacobi_omega;
        if (cfg->amg_truncation >= 0)
            p.truncation = cfg->amg_truncation;
    }
    auto H = std::make_unique<mi_amg_host>();
    if (amg_build(A, p, H->h) != 0)
        return set_error(MI_ERR_BAD_ARG, "hierarchy construction failed");
    *out = H.release();
    return MI_OK;
}

MI_API int mi_amg_host_levels(const mi_amg_host* H, int* levels, double* operator_complexity,
                              double* grid_complexity)
{
    if (!H)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    if (levels)
        *levels = int(H->h.levels.size());
    if (operator_complexity)
        *operator_complexity = H->h.operator_complexity;
    if (grid_complexity)
        *grid_complexity = H->h.grid_complexity;
    return MI_OK;
}

// which: 0 = A_l, 1 = P_l (level l+1 -> l), 2 = R_l = P_l^T, 3 = coarsest dense inverse,
// 4 = aggregates of the rows of A_l, 5 = strong-coupling mask of A_l's entries
MI_API int mi_amg_host_level_info(const mi_amg_host* H, int level, int which, int64_t* nrows,
                                  int64_t* ncols, int64_t* nnz, double* lambda_max)
{
    if (!H || level < 0 || level >= int(H->h.levels.size()) || which < 0 || which > 5)
        return set_error(MI_ERR_BAD_ARG, "bad level or matrix selector");
    const AmgLevelHost& L = H->h.levels[size_t(level)];
    if (which >= 4) {
        if (L.agg.empty() || L.strong.empty())
            return set_error(MI_ERR_UNSUPPORTED, "level %d records no aggregates", level);
        if (nrows)
            *nrows = L.A.nr;
        if (ncols)
            *ncols = which == 4 ? L.n_aggregates : L.A.nc;
        if (nnz)
            *nnz = which == 4 ? int64_t(L.agg.size()) : int64_t(L.strong.size());
    } else if (which == 3) {
        const int64_t c = H->h.coarse_n;
        if (nrows)
            *nrows = c;
        if (ncols)
            *ncols = c;
        if (nnz)
            *nnz = c * c;
    } else {
        const HCsr& M = which == 0 ? L.A : which == 1 ? L.P : L.R;
        if (nrows)
            *nrows = M.nr;
        if (ncols)
            *ncols = M.nc;
        if (nnz)
            *nnz = M.nnz();
    }
    if (lambda_max)
        *lambda_max = L.lmax;
    return MI_OK;
}

MI_API int mi_amg_host_level_get(const mi_amg_host* H, int level, int which, int32_t* rp,
                                 int32_t* ci, double* v)
{
    if (!H || level < 0 || level >= int(H->h.levels.size()) || which < 0 || which > 5)
        return set_error(MI_ERR_BAD_ARG, "bad level or matrix selector");
    if (which >= 4) {
        const AmgLevelHost& L = H->h.levels[size_t(level)];
        if (L.agg.empty() || L.strong.empty())
            return set_error(MI_ERR_UNSUPPORTED, "level %d records no aggregates", level);
        if (ci && which == 4)
            std::memcpy(ci, L.agg.data(), L.agg.size() * sizeof(int32_t));
        if (ci && which == 5)
            for (size_t k = 0; k < L.strong.size(); ++k)
                ci[k] = L.strong[k];
        return MI_OK;
    }
    if (which == 3) {
        if (v)
            std::memcpy(v, H->h.coarse_inverse.data(), H->h.coarse_inverse.size() * sizeof(double));
        return MI_OK;
    }
    const AmgLevelHost& L = H->h.levels[size_t(level)];
    const HCsr& M = which == 0 ? L.A : which == 1 ? L.P : L.R;
    if (rp)
        std::memcpy(rp, M.rp.data(), M.rp.size() * sizeof(int32_t));
    if (ci)
        std::memcpy(ci, M.ci.data(), M.ci.size() * sizeof(int32_t));
    if (v)
        std::memcpy(v, M.v.data(), M.v.size() * sizeof(double));
    return MI_OK;
}

MI_API int mi_amg_host_destroy(mi_amg_host* H)
{
    delete H;
    return MI_OK;
}

MI_API int mi_rcm_permutation(int64_t n, const int32_t* rp, const int32_t* ci, int32_t* perm,
                              int32_t* bandwidth_after)
{
    if (n < 0 || n > INT32_MAX - 1 || (n > 0 && (!rp || !ci || !perm)))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    MI_TRY(check_host_csr(n, rp, ci));
    std::vector<int> p;
    rcm_order(int(n), rp, ci, p);
    std::memcpy(perm, p.data(), p.size() * sizeof(int32_t));
    if (bandwidth_after) {
        std::vector<int> brp, bci, vmap;
        permute_pattern(int(n), rp, ci, p, brp, bci, vmap);
        *bandwidth_after = bandwidth(int(n), brp.data(), bci.data());
    }
    return MI_OK;
}

MI_API int mi_color_permutation(int64_t n, const int32_t* rp, const int32_t* ci, int32_t* perm,
                                int32_t* colors)
{
    if (n < 0 || n > INT32_MAX - 1 || (n > 0 && (!rp || !ci || !perm)))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    MI_TRY(check_host_csr(n, rp, ci));
    std::vector<int> p;
    int nc = 0;
    color_order(int(n), rp, ci, p, &nc);
    std::memcpy(perm, p.data(), p.size() * sizeof(int32_t));
    if (colors)
        *colors = nc;
    return MI_OK;
}

} // extern "C"
