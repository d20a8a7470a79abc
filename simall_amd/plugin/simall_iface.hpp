// simall_iface.hpp — the SimAll interfaces this backend plugs into, declared
// with the same names, namespaces, member layout and virtual-function order
// as the host application's headers so that objects cross the boundary:
//
//   simall::util::AlignedAllocator / aligned_vector  (src/utilities/AlignedAllocator.hpp:57-102)
//   simall::solver::CSRMatrix                        (src/solver/CSRMatrix.hpp:16-36)
//   LinearSolverKind / PreconditionerKind /
//   LinearSolverConfig / ILinearSolver                (src/solver/Solver.hpp:67-100)
//   simall::plugins::IPlugin, kPluginAbiVersion       (plugins/IPlugin.hpp:20-42)
//   PluginBase, PluginCategory, CategorisedPlugin     (plugins/SimAllPluginSdk.hpp:36-106, 270-280)
//
// Building inside the SimAll tree, define SIMALL_HIP_USE_HOST_HEADERS and put
// its src/ and plugins/ directories on the include path to use the host's
// own headers instead.
#pragma once

#if defined(SIMALL_HIP_USE_HOST_HEADERS)

#include "SimAllPluginSdk.hpp"
#include "solver/LinearSolvers.hpp"
#include "solver/Solver.hpp"

#else

#include <cstddef>
#include <cstdlib>
#include <limits>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

namespace simall::util
{

template <typename T, std::size_t Alignment = 64> class AlignedAllocator
{
public:
    using value_type = T;
    using size_type = std::size_t;
    using difference_type = std::ptrdiff_t;
    using propagate_on_container_move_assignment = std::true_type;
    using is_always_equal = std::true_type;
    template <typename U> struct rebind
    {
        using other = AlignedAllocator<U, Alignment>;
    };
    constexpr AlignedAllocator() noexcept = default;
    template <typename U> constexpr AlignedAllocator(const AlignedAllocator<U, Alignment>&) noexcept {}
    [[nodiscard]] T* allocate(size_type n)
    {
        if (n > std::numeric_limits<size_type>::max() / sizeof(T))
            throw std::bad_array_new_length{};
        if (n == 0)
            return nullptr;
        void* p = nullptr;
        if (posix_memalign(&p, Alignment, n * sizeof(T)) != 0)
            throw std::bad_alloc{};
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_type) noexcept { std::free(p); }
    template <typename U, std::size_t A2>
    constexpr bool operator==(const AlignedAllocator<U, A2>&) const noexcept
    {
        return Alignment == A2;
    }
    template <typename U, std::size_t A2>
    constexpr bool operator!=(const AlignedAllocator<U, A2>&) const noexcept
    {
        return Alignment != A2;
    }
};

template <typename T> using aligned_vector = std::vector<T, AlignedAllocator<T>>;

} // namespace simall::util

namespace simall::solver
{

struct CSRMatrix
{
    util::aligned_vector<int> rowPtr;
    util::aligned_vector<int> colIdx;
    util::aligned_vector<double> values;
    std::size_t rows() const noexcept { return rowPtr.empty() ? 0 : rowPtr.size() - 1; }
};

enum class LinearSolverKind
{
    GMRES,
    BiCGSTAB,
    CG,
    TFQMR
};
enum class PreconditionerKind
{
    Jacobi,
    ILU,
    SSOR,
    AMG,
    None
};

struct LinearSolverConfig
{
    LinearSolverKind kind = LinearSolverKind::GMRES;
    PreconditionerKind preconditioner = PreconditionerKind::ILU;
    double tolerance = 1.0e-8;
    int maxIterations = 1000;
    int restart = 30;
};

class ILinearSolver
{
public:
    virtual ~ILinearSolver() = default;
    virtual int solve(const CSRMatrix& A, const util::aligned_vector<double>& b,
                      util::aligned_vector<double>& x) = 0;
    virtual double last_residual() const = 0;
};

} // namespace simall::solver

namespace simall::plugins
{

inline constexpr int kPluginAbiVersion = 1;

class IPlugin
{
public:
    virtual ~IPlugin() = default;
    virtual std::string name() const = 0;
    virtual std::string version() const = 0;
    virtual void on_load() = 0;
    virtual void on_unload() = 0;
    virtual int abi() const { return kPluginAbiVersion; }
};

class PluginBase : public IPlugin
{
public:
    PluginBase() = default;
    ~PluginBase() override = default;
    PluginBase(const PluginBase&) = delete;
    PluginBase& operator=(const PluginBase&) = delete;
    std::string name() const override { return "unnamed"; }
    std::string version() const override { return "0.0.0"; }
    void on_load() override {}
    void on_unload() override {}
    int abi() const override { return kPluginAbiVersion; }
};

// Only the ordinals this plugin uses; values are part of the plugin ABI.
enum class PluginCategory
{
    Solver = 0,
    Gpu = 23,
};

class CategorisedPlugin : public PluginBase
{
public:
    explicit CategorisedPlugin(PluginCategory cat) : category_(cat) {}
    [[nodiscard]] PluginCategory category() const { return category_; }

private:
    PluginCategory category_;
};

} // namespace simall::plugins

#if defined(_WIN32)
#define SIMALL_PLUGIN_EXPORT __declspec(dllexport)
#else
#define SIMALL_PLUGIN_EXPORT __attribute__((visibility("default")))
#endif

#endif
