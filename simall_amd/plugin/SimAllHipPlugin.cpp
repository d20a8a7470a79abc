// SimAllHipPlugin.cpp — the backend as a SimAll plugin module.
//
// Exports the plugin ABI (plugins/IPlugin.hpp:41-42; ABI version 1) so that
// core::PluginLoader (src/core/PluginLoader.cpp:68-163) can load it, plus a
// flat factory.  make_linear_solver is a closed switch
// (src/solver/Solver.cpp:25-39); the host opts in by exporting
//   extern "C" void simall_register_linear_solver_factory(
//       simall::solver::ILinearSolver* (*)(const simall::solver::LinearSolverConfig*));
// and consulting the registered factory first.  on_load() registers with it
// when the symbol exists and a gfx950 device is present; otherwise it only
// reports unavailability (it never throws).
#include "HipLinearSolver.hpp"

#include <dlfcn.h>

#include <cstdio>

namespace
{

using Factory = simall::solver::ILinearSolver* (*)(const simall::solver::LinearSolverConfig*);
using RegisterFn = void (*)(Factory);

class SimAllHipPlugin final : public simall::plugins::CategorisedPlugin
{
public:
    SimAllHipPlugin() : CategorisedPlugin(simall::plugins::PluginCategory::Solver) {}
    std::string name() const override { return "simall_hip_solver"; }
    std::string version() const override { return "1.0.0"; }
    void on_load() override
    {
        available_ = mi_init(0) == MI_OK;
        if (!available_) {
            std::fprintf(stderr, "[plugin] simall_hip_solver: no gfx950 device (%s); not registered\n",
                         mi_last_error());
            return;
        }
        auto reg = reinterpret_cast<RegisterFn>(dlsym(RTLD_DEFAULT, "simall_register_linear_solver_factory"));
        if (reg) {
            reg(&create);
            registered_ = true;
        }
    }
    void on_unload() override
    {
        auto reg = reinterpret_cast<RegisterFn>(dlsym(RTLD_DEFAULT, "simall_register_linear_solver_factory"));
        if (registered_ && reg)
            reg(nullptr);
        registered_ = false;
        if (available_)
            mi_device_synchronize();
    }
    bool available() const { return available_; }

    static simall::solver::ILinearSolver* create(const simall::solver::LinearSolverConfig* cfg)
    {
        if (!cfg)
            return nullptr;
        return simall::hip::HipLinearSolver::create(*cfg).release();
    }

private:
    bool available_ = false;
    bool registered_ = false;
};

} // namespace

extern "C" SIMALL_PLUGIN_EXPORT simall::plugins::IPlugin* CreatePlugin()
{
    return new SimAllHipPlugin();
}

extern "C" SIMALL_PLUGIN_EXPORT int GetPluginAbi() { return simall::plugins::kPluginAbiVersion; }

// Flat factory for hosts without the registry hook: a new ILinearSolver for
// cfg, or nullptr when the configuration has no device path.
extern "C" SIMALL_PLUGIN_EXPORT simall::solver::ILinearSolver*
SimAllHip_CreateLinearSolver(const simall::solver::LinearSolverConfig* cfg, int precision)
{
    if (!cfg)
        return nullptr;
    return simall::hip::HipLinearSolver::create(*cfg, precision).release();
}
