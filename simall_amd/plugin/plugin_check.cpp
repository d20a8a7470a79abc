// plugin_check.cpp — loads the plugin module the way core::PluginLoader does
// (src/core/PluginLoader.cpp:68-163: dlopen RTLD_NOW|RTLD_LOCAL, ABI check,
// CreatePlugin, on_load inside try/catch, on_unload, dlclose) and, in "solve"
// mode, drives the device solver through solver::ILinearSolver.
//
//   plugin_check <module.so> load    ABI / identity / no-throw lifecycle
//   plugin_check <module.so> solve   + factory registration and solves
//
// Prints one line per check and "OK" at the end; exits non-zero on failure.
#include "simall_iface.hpp"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

using namespace simall;

namespace
{
solver::ILinearSolver* (*g_factory)(const solver::LinearSolverConfig*) = nullptr;
int g_registrations = 0;

int fail(const char* what)
{
    std::printf("FAIL: %s\n", what);
    return 1;
}

solver::CSRMatrix tridiag(int n, double lo, double d, double up)
{
    solver::CSRMatrix A;
    A.rowPtr.push_back(0);
    for (int i = 0; i < n; ++i) {
        if (i > 0) {
            A.colIdx.push_back(i - 1);
            A.values.push_back(lo);
        }
        A.colIdx.push_back(i);
        A.values.push_back(d);
        if (i + 1 < n) {
            A.colIdx.push_back(i + 1);
            A.values.push_back(up);
        }
        A.rowPtr.push_back(int(A.colIdx.size()));
    }
    return A;
}

double residual(const solver::CSRMatrix& A, const util::aligned_vector<double>& b,
                const util::aligned_vector<double>& x)
{
    double s2 = 0.0;
    for (std::size_t i = 0; i < A.rows(); ++i) {
        double s = 0.0;
        for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; ++k)
            s += A.values[size_t(k)] * x[size_t(A.colIdx[size_t(k)])];
        s2 += (b[i] - s) * (b[i] - s);
    }
    return std::sqrt(s2);
}
} // namespace

// The registry hook a host application exports (see SimAllHipPlugin.cpp).
extern "C" __attribute__((visibility("default"))) void
simall_register_linear_solver_factory(solver::ILinearSolver* (*f)(const solver::LinearSolverConfig*))
{
    g_factory = f;
    ++g_registrations;
}

int main(int argc, char** argv)
{
    if (argc < 3)
        return fail("usage: plugin_check <module.so> load|solve");
    const bool solve = std::strcmp(argv[2], "solve") == 0;
    void* h = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        std::printf("dlopen: %s\n", dlerror());
        return fail("dlopen");
    }
    auto abi = reinterpret_cast<int (*)()>(dlsym(h, "GetPluginAbi"));
    auto create = reinterpret_cast<plugins::IPlugin* (*)()>(dlsym(h, "CreatePlugin"));
    if (!abi || !create)
        return fail("missing GetPluginAbi / CreatePlugin");
    if (abi() != plugins::kPluginAbiVersion)
        return fail("ABI version mismatch");
    std::unique_ptr<plugins::IPlugin> p(create());
    if (!p || p->abi() != 1)
        return fail("CreatePlugin");
    auto* cat = dynamic_cast<plugins::CategorisedPlugin*>(p.get());
    if (!cat || cat->category() != plugins::PluginCategory::Solver)
        return fail("category");
    std::printf("plugin %s %s abi %d category %d\n", p->name().c_str(), p->version().c_str(),
                p->abi(), int(cat->category()));
    try {
        p->on_load();
    } catch (...) {
        return fail("on_load threw");
    }
    std::printf("registered factory: %s\n", g_factory ? "yes" : "no");
    if (solve) {
        if (!g_factory)
            return fail("on_load did not register the factory");
        solver::LinearSolverConfig cfg;
        cfg.kind = solver::LinearSolverKind::TFQMR; // no device path: host keeps its own
        if (g_factory(&cfg) != nullptr)
            return fail("TFQMR should not be claimed");
        {
            // tests/unit/solver/test_ssor_preconditioner.cpp:65-88 through ILinearSolver
            solver::LinearSolverConfig c2;
            c2.kind = solver::LinearSolverKind::CG;
            c2.preconditioner = solver::PreconditionerKind::SSOR;
            c2.maxIterations = 200;
            c2.tolerance = 1e-10;
            std::unique_ptr<solver::ILinearSolver> ss(g_factory(&c2));
            if (!ss)
                return fail("CG/SSOR factory");
            const auto L50 = tridiag(50, -1.0, 2.0, -1.0);
            util::aligned_vector<double> b50(50, 1.0), x50;
            const int its = ss->solve(L50, b50, x50);
            const double rs = residual(L50, b50, x50);
            std::printf("CG/SSOR n=50: %d iterations, ||r|| %.3e\n", its, rs);
            if (!(rs < 1e-7 && its > 0))
                return fail("CG/SSOR solve through ILinearSolver");
        }
        {
            // GMRES(20) + ILU(0) as the turbulence models configure it
            // (src/turbulence/KOmegaSST.cpp:32-37), on a nonsymmetric system
            solver::LinearSolverConfig g;
            g.kind = solver::LinearSolverKind::GMRES;
            g.preconditioner = solver::PreconditionerKind::ILU;
            g.tolerance = 1e-6;
            g.maxIterations = 100;
            g.restart = 20;
            std::unique_ptr<solver::ILinearSolver> sg(g_factory(&g));
            if (!sg)
                return fail("GMRES/ILU factory");
            const auto N = tridiag(500, -1.3, 2.5, -0.7);
            util::aligned_vector<double> bg(500, 1.0), xg;
            const int itg = sg->solve(N, bg, xg);
            const double rg = residual(N, bg, xg) / std::sqrt(500.0);
            std::printf("GMRES(20)/ILU n=500: %d iterations, rel %.2e, last_residual %.3e\n", itg,
                        rg, sg->last_residual());
            if (!(itg < 100 && rg < 1e-5 && sg->last_residual() < 1e-6))
                return fail("GMRES solve through ILinearSolver");
            g.kind = solver::LinearSolverKind::BiCGSTAB;
            std::unique_ptr<solver::ILinearSolver> sb(g_factory(&g));
            util::aligned_vector<double> xb;
            const int itb = sb ? sb->solve(N, bg, xb) : -1;
            const double rb = residual(N, bg, xb) / std::sqrt(500.0);
            std::printf("BiCGStab/ILU n=500: %d iterations, rel %.2e\n", itb, rb);
            if (!(itb >= 0 && itb < 100 && rb < 1e-5))
                return fail("BiCGStab solve through ILinearSolver");
        }
        // tests/unit/gpu/test_gpu_models.cpp:181-235 through ILinearSolver
        cfg.kind = solver::LinearSolverKind::CG;
        cfg.preconditioner = solver::PreconditionerKind::Jacobi;
        cfg.tolerance = 1e-10;
        cfg.maxIterations = 200;
        std::unique_ptr<solver::ILinearSolver> s(g_factory(&cfg));
        if (!s)
            return fail("CG/Jacobi factory");
        const auto A = tridiag(10, -1.0, 4.0, -1.0);
        util::aligned_vector<double> b(10, 0.0), x(10, 123.0);
        b.front() = b.back() = 1.0;
        const int it = s->solve(A, b, x);
        const double r = residual(A, b, x);
        std::printf("CG/Jacobi n=10: %d iterations, ||r|| %.3e, last_residual %.3e\n", it, r,
                    s->last_residual());
        if (!(it < 50 && r < 1e-8 && s->last_residual() < 1e-10))
            return fail("tridiagonal solve");
        // iteration limit: returns maxIterations, residual above tolerance
        cfg.maxIterations = 3;
        cfg.tolerance = 1e-14;
        std::unique_ptr<solver::ILinearSolver> s3(g_factory(&cfg));
        const auto L = tridiag(100, -1.0, 2.0, -1.0);
        util::aligned_vector<double> b3(100, 1.0), x3;
        const int it3 = s3->solve(L, b3, x3);
        std::printf("iteration limit: returned %d, last_residual %.3e\n", it3, s3->last_residual());
        if (it3 != 3 || !(s3->last_residual() > 1e-14) || x3.size() != 100)
            return fail("iteration-limit semantics");
        // AMG preconditioner, new values on the same pattern
        cfg.preconditioner = solver::PreconditionerKind::AMG;
        cfg.maxIterations = 200;
        cfg.tolerance = 1e-8;
        std::unique_ptr<solver::ILinearSolver> sa(g_factory(&cfg));
        auto L2 = tridiag(2000, -1.0, 2.001, -1.0);
        util::aligned_vector<double> b4(2000, 1.0), x4;
        int ita = sa->solve(L2, b4, x4);
        const double r4 = residual(L2, b4, x4) / std::sqrt(2000.0);
        for (auto& v : L2.values)
            v *= 2.0;
        int itb = sa->solve(L2, b4, x4);
        const double r5 = residual(L2, b4, x4) / std::sqrt(2000.0);
        std::printf("CG/AMG n=2000: %d and %d iterations, rel %.2e %.2e\n", ita, itb, r4, r5);
        if (!(r4 < 1e-8 && r5 < 1e-8 && ita < 200 && itb < 200))
            return fail("AMG solve through ILinearSolver");
    }
    try {
        p->on_unload();
    } catch (...) {
        return fail("on_unload threw");
    }
    p.reset();
    dlclose(h);
    std::printf("OK\n");
    return 0;
}
