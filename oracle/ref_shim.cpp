// ref_shim.cpp — flat C entry points over the reference's own sources, built
// only in a container where /root/reference exists (oracle/ref.mk), into
// oracle/_ref/libsimall_ref.so.  TEST INFRASTRUCTURE: used by tests/ to pin
// the oracle and the workload generator against the real code; never part of
// the product.
//
// Reference code called (not copied): meshing::CartesianMesher / Mesh
// (src/meshing/CartesianMesher.cpp, Connectivity.cpp, MeshStorage.cpp),
// solver::SimpleAlgorithm assembly (src/solver/SimpleAlgorithm.cpp, through a
// subclass that reaches its protected hooks), solver::CSRMatrix::spmv,
// adjoint::solve_discrete_adjoint (src/adjoint/DiscreteAdjoint.cpp) and the
// Ruge-Stueben AMGPreconditioner (src/solver/AMGPreconditioner.cpp),
// LeastSquaresGradient (src/solver/Gradient.cpp) and the SIMPLE / PISO outer
// iterations (SimpleAlgorithm.cpp, PisoAlgorithm.cpp) with a dense direct
// solve standing in for the linear solvers, and the SIMALLCP checkpoint
// writer and reader (src/io/Checkpoint.cpp).
#include "adjoint/DiscreteAdjoint.hpp"
#include "io/Checkpoint.hpp"
#include "meshing/CartesianMesher.hpp"
#include "solver/AMGPreconditioner.hpp"
#include "solver/Gradient.hpp"
#include "solver/PisoAlgorithm.hpp"
#include "solver/SimpleAlgorithm.hpp"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>

using namespace simall;

namespace
{

uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return double(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// A mesh from the reference's generator; interior nodes optionally moved by
// the same seeded sequence as the build's workload generator, then the
// reference's geometry pass.
std::unique_ptr<meshing::Mesh> make_mesh(int Nx, int Ny, int Nz, const double* o, const double* L,
                                         double perturb, uint64_t seed)
{
    meshing::CartesianGridSpec spec;
    spec.origin = util::Vec3d{o[0], o[1], o[2]};
    spec.extent = util::Vec3d{L[0], L[1], L[2]};
    spec.Nx = Nx;
    spec.Ny = Ny;
    spec.Nz = Nz;
    auto m = std::make_unique<meshing::Mesh>();
    meshing::CartesianMesher(spec).generate(*m);
    if (perturb > 0.0) {
        const double dx = L[0] / Nx, dy = L[1] / Ny, dz = L[2] / Nz;
        auto& N = m->nodes();
        uint64_t s = seed;
        const int Px = Nx + 1, Py = Ny + 1, Pz = Nz + 1;
        for (int k = 0; k < Pz; ++k)
            for (int j = 0; j < Py; ++j)
                for (int i = 0; i < Px; ++i) {
                    if (!(i > 0 && i < Nx && j > 0 && j < Ny && k > 0 && k < Nz))
                        continue;
                    const size_t id = size_t((k * Py + j) * Px + i);
                    N.x[id] += (2.0 * uniform01(s) - 1.0) * perturb * dx;
                    N.y[id] += (2.0 * uniform01(s) - 1.0) * perturb * dy;
                    N.z[id] += (2.0 * uniform01(s) - 1.0) * perturb * dz;
                }
        m->compute_geometry();
    }
    return m;
}

class NullSolver final : public solver::ILinearSolver
{
public:
    int solve(const solver::CSRMatrix&, const util::aligned_vector<double>&,
              util::aligned_vector<double>&) override
    {
        return 0;
    }
    double last_residual() const override { return 0.0; }
};

// Reaches SimpleAlgorithm's protected assembly hooks and storage.
class Probe final : public solver::SimpleAlgorithm
{
public:
    using SimpleAlgorithm::SimpleAlgorithm;
    void sparsity() { build_sparsity(); }
    void pressure(const double* aP, const double* phi, const double* amom = nullptr)
    {
        std::memcpy(aP_.data(), aP, aP_.size() * sizeof(double));
        if (amom)
            std::memcpy(A_mom_.values.data(), amom, A_mom_.values.size() * sizeof(double));
        phiFlux_.assign(phi, phi + mesh_.faces().size());
        assemble_pressure_correction();
    }
    void momentum(int comp, const double* phi, const double* Uold)
    {
        phiFlux_.assign(phi, phi + mesh_.faces().size());
        auto& Un = comp == 0 ? Ux_n_ : comp == 1 ? Uy_n_ : Uz_n_;
        if (Uold)
            std::memcpy(Un.data(), Uold, Un.size() * sizeof(double));
        assemble_momentum(comp);
    }
    void flux() { compute_face_fluxes(); }
    void correct(const double* aP, const double* pp, const double* phi)
    {
        std::memcpy(aP_.data(), aP, aP_.size() * sizeof(double));
        std::memcpy(pPrime_.data(), pp, pPrime_.size() * sizeof(double));
        phiFlux_.assign(phi, phi + mesh_.faces().size());
        correct_fields();
    }
    const solver::CSRMatrix& Ap() const { return A_p_; }
    const solver::CSRMatrix& Am() const { return A_mom_; }
    const std::vector<int>& twins() const { return periodicTwin_; }
    const util::aligned_vector<double>& rhs() const { return rhs_; }
    const util::aligned_vector<double>& phi() const { return phiFlux_; }
};

// A dense LU solve with partial pivoting as the ILinearSolver of the outer
// iterations: their results then depend on the assembly and correction only.
class DenseSolver final : public solver::ILinearSolver
{
public:
    int solve(const solver::CSRMatrix& A, const util::aligned_vector<double>& b,
              util::aligned_vector<double>& x) override
    {
        const size_t n = A.rows();
        std::vector<double> M(n * n, 0.0);
        for (size_t i = 0; i < n; ++i)
            for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; ++k)
                M[i * n + size_t(A.colIdx[size_t(k)])] += A.values[size_t(k)];
        std::vector<double> y(b.begin(), b.end());
        for (size_t c = 0; c < n; ++c) {
            size_t piv = c;
            for (size_t r = c + 1; r < n; ++r)
                if (std::fabs(M[r * n + c]) > std::fabs(M[piv * n + c]))
                    piv = r;
            if (piv != c) {
                for (size_t k = 0; k < n; ++k)
                    std::swap(M[c * n + k], M[piv * n + k]);
                std::swap(y[c], y[piv]);
            }
            const double d = M[c * n + c];
            for (size_t r = c + 1; r < n; ++r) {
                const double f = M[r * n + c] / d;
                if (f == 0.0)
                    continue;
                for (size_t k = c; k < n; ++k)
                    M[r * n + k] -= f * M[c * n + k];
                y[r] -= f * y[c];
            }
        }
        x.assign(n, 0.0);
        for (size_t i = n; i-- > 0;) {
            double s = y[i];
            for (size_t k = i + 1; k < n; ++k)
                s -= M[i * n + k] * x[k];
            x[i] = s / M[i * n + i];
        }
        return 1;
    }
    double last_residual() const override { return 0.0; }
};

// PisoAlgorithm with access to its face flux.
class PisoProbe final : public solver::PisoAlgorithm
{
public:
    using PisoAlgorithm::PisoAlgorithm;
    const util::aligned_vector<double>& phi() const { return phiFlux_; }
};

class SimpleProbe final : public solver::SimpleAlgorithm
{
public:
    using SimpleAlgorithm::SimpleAlgorithm;
    const util::aligned_vector<double>& phi() const { return phiFlux_; }
};

solver::BCType bc_type(int cls)
{
    switch (cls) {
    case 0: return solver::BCType::Wall;
    case 1: return solver::BCType::PressureOutlet;
    case 2: return solver::BCType::Symmetry;
    case 3: return solver::BCType::VelocityInlet;
    case 5: return solver::BCType::Periodic;
    default: return solver::BCType::OutFlow;
    }
}

struct Case
{
    std::unique_ptr<meshing::Mesh> mesh;
    solver::FieldRegistry fields;
    std::vector<solver::BoundarySpec> bcs;
    NullSolver s1, s2;
    std::unique_ptr<Probe> probe;
};

std::unique_ptr<Case> make_case(int Nx, int Ny, int Nz, const double* o, const double* L,
                                double perturb, uint64_t seed, const int* bc_of_zone, int nz,
                                const double* bc_value, double rho, double mu, double dt, double urf,
                                const double* U, bool simplec = false)
{
    auto c = std::make_unique<Case>();
    c->mesh = make_mesh(Nx, Ny, Nz, o, L, perturb, seed);
    const size_t nC = c->mesh->cells().size();
    for (int z = 1; z < nz; ++z) {
        solver::BoundarySpec b;
        b.zone = meshing::ZoneId(z);
        b.type = bc_type(bc_of_zone[z]);
        for (int d = 0; d < 3; ++d)
            b.vectorValue[d] = bc_value ? bc_value[z * 3 + d] : 0.0;
        c->bcs.push_back(b);
    }
    auto& Uf = c->fields.vector("U", nC);
    c->fields.scalar("p", nC);
    if (U) {
        std::memcpy(Uf.x.data(), U, nC * sizeof(double));
        std::memcpy(Uf.y.data(), U + nC, nC * sizeof(double));
        std::memcpy(Uf.z.data(), U + 2 * nC, nC * sizeof(double));
    }
    solver::SimpleOptions opt;
    opt.rho = rho;
    opt.mu = mu;
    opt.dt = dt;
    opt.urfU = urf;
    opt.timeScheme = solver::TemporalScheme::ImplicitEuler;
    if (simplec)
        opt.algorithm = solver::PvCouplingVariant::SIMPLEC;
    c->probe = std::make_unique<Probe>(*c->mesh, c->fields, c->bcs, c->s1, c->s2, opt);
    c->probe->sparsity();
    return c;
}

void copy_csr(const solver::CSRMatrix& A, int32_t* rp, int32_t* ci, double* v)
{
    if (rp)
        std::memcpy(rp, A.rowPtr.data(), A.rowPtr.size() * sizeof(int32_t));
    if (ci)
        std::memcpy(ci, A.colIdx.data(), A.colIdx.size() * sizeof(int32_t));
    if (v)
        std::memcpy(v, A.values.data(), A.values.size() * sizeof(double));
}

} // namespace

extern "C" {

// Sizes of the reference mesh: cells, faces, nnz of the SIMPLE pattern.
int ref_mesh_sizes(int Nx, int Ny, int Nz, const double* o, const double* L, int64_t* nC,
                   int64_t* nF, int64_t* nnz)
{
    const int bc[7] = {0, 0, 0, 0, 0, 0, 0};
    auto c = make_case(Nx, Ny, Nz, o, L, 0.0, 0, bc, 7, nullptr, 1, 1, 0, 1, nullptr);
    *nC = int64_t(c->mesh->cells().size());
    *nF = int64_t(c->mesh->faces().size());
    *nnz = int64_t(c->probe->Ap().colIdx.size());
    return 0;
}

// Mesh arrays after the reference generator (+ optional perturbation and the
// reference geometry pass).
int ref_mesh_arrays(int Nx, int Ny, int Nz, const double* o, const double* L, double perturb,
                    uint64_t seed, uint64_t* owner, uint64_t* neighbor, double* area, double* fc,
                    uint32_t* zone, double* vol, double* cc, int32_t* cfo, uint64_t* cfi)
{
    auto m = make_mesh(Nx, Ny, Nz, o, L, perturb, seed);
    const auto& F = m->faces();
    const auto& C = m->cells();
    const size_t nF = F.size(), nC = C.size();
    std::memcpy(owner, F.owner.data(), nF * 8);
    std::memcpy(neighbor, F.neighbor.data(), nF * 8);
    std::memcpy(area, F.areaX.data(), nF * 8);
    std::memcpy(area + nF, F.areaY.data(), nF * 8);
    std::memcpy(area + 2 * nF, F.areaZ.data(), nF * 8);
    std::memcpy(fc, F.centroidX.data(), nF * 8);
    std::memcpy(fc + nF, F.centroidY.data(), nF * 8);
    std::memcpy(fc + 2 * nF, F.centroidZ.data(), nF * 8);
    std::memcpy(zone, F.boundaryZone.data(), nF * 4);
    std::memcpy(vol, C.volume.data(), nC * 8);
    std::memcpy(cc, C.centroidX.data(), nC * 8);
    std::memcpy(cc + nC, C.centroidY.data(), nC * 8);
    std::memcpy(cc + 2 * nC, C.centroidZ.data(), nC * 8);
    std::memcpy(cfo, C.faceOffsets.data(), (nC + 1) * 4);
    std::memcpy(cfi, C.faceIndices.data(), size_t(C.faceOffsets[nC]) * 8);
    return 0;
}

// SimpleAlgorithm::assemble_pressure_correction on the reference mesh.
int ref_assemble_pressure(int Nx, int Ny, int Nz, const double* o, const double* L, double perturb,
                          uint64_t seed, const int* bc_of_zone, int nz, double rho,
                          const double* aP, const double* phi, int32_t* rp, int32_t* ci,
                          double* vals, double* rhs)
{
    auto c = make_case(Nx, Ny, Nz, o, L, perturb, seed, bc_of_zone, nz, nullptr, rho, 1e-3, 0.0,
                       1.0, nullptr);
    c->probe->pressure(aP, phi);
    copy_csr(c->probe->Ap(), rp, ci, vals);
    std::memcpy(rhs, c->probe->rhs().data(), c->probe->rhs().size() * sizeof(double));
    return 0;
}

// The periodic face pairs (twin[n_faces]), the SIMPLE pattern with the twin
// columns (rp, ci; nnz returned when ci is null) of build_sparsity.
int64_t ref_periodic(int Nx, int Ny, int Nz, const double* o, const double* L, double perturb,
                     uint64_t seed, const int* bc_of_zone, int nz, int32_t* twin, int32_t* rp,
                     int32_t* ci)
{
    auto c = make_case(Nx, Ny, Nz, o, L, perturb, seed, bc_of_zone, nz, nullptr, 1.0, 1e-3, 0.0,
                       1.0, nullptr);
    const auto& tw = c->probe->twins();
    if (twin)
        for (size_t f = 0; f < c->mesh->faces().size(); ++f)
            twin[f] = tw.empty() ? -1 : tw[f];
    copy_csr(c->probe->Ap(), rp, ci, nullptr);
    return int64_t(c->probe->Ap().colIdx.size());
}

// assemble_pressure_correction with the SIMPLEC denominators: the momentum
// matrix values amom (on the pattern) and aP.
int ref_assemble_pressure_simplec(int Nx, int Ny, int Nz, const double* o, const double* L,
                                  double perturb, uint64_t seed, const int* bc_of_zone, int nz,
                                  double rho, const double* aP, const double* amom,
                                  const double* phi, double* vals, double* rhs)
{
    auto c = make_case(Nx, Ny, Nz, o, L, perturb, seed, bc_of_zone, nz, nullptr, rho, 1e-3, 0.0,
                       1.0, nullptr, true);
    c->probe->pressure(aP, phi, amom);
    copy_csr(c->probe->Ap(), nullptr, nullptr, vals);
    std::memcpy(rhs, c->probe->rhs().data(), c->probe->rhs().size() * sizeof(double));
    return 0;
}

// compute_face_fluxes + assemble_momentum(comp) (p = 0) on the reference mesh.
int ref_assemble_momentum(int Nx, int Ny, int Nz, const double* o, const double* L, double perturb,
                          uint64_t seed, const int* bc_of_zone, int nz, const double* bc_value,
                          double rho, double mu, double dt, double urf, const double* U, int comp,
                          const double* Uold, double* phi, double* vals, double* rhs)
{
    auto c = make_case(Nx, Ny, Nz, o, L, perturb, seed, bc_of_zone, nz, bc_value, rho, mu, dt, urf,
                       U);
    c->probe->flux();
    const size_t nF = c->mesh->faces().size();
    std::memcpy(phi, c->probe->phi().data(), nF * sizeof(double));
    c->probe->momentum(comp, phi, Uold);
    copy_csr(c->probe->Am(), nullptr, nullptr, vals);
    std::memcpy(rhs, c->probe->rhs().data(), c->probe->rhs().size() * sizeof(double));
    return 0;
}

// solver::CSRMatrix::spmv
int ref_spmv(int n, const int32_t* rp, const int32_t* ci, const double* v, const double* x, double* y)
{
    solver::CSRMatrix A;
    A.rowPtr.assign(rp, rp + n + 1);
    A.colIdx.assign(ci, ci + rp[n]);
    A.values.assign(v, v + rp[n]);
    util::aligned_vector<double> X(x, x + n), Y;
    A.spmv(X, Y);
    std::memcpy(y, Y.data(), size_t(n) * sizeof(double));
    return 0;
}

// adjoint::solve_discrete_adjoint with (dR/dalpha)^T psi supplied as a
// dense m x n row-major matrix G: out = G psi.
int ref_discrete_adjoint(int n, const int32_t* rp, const int32_t* ci, const double* v,
                         const double* dJdQ, int m, const double* dJdAlpha, const double* G,
                         int maxIter, int restart, double tol, double* psi, double* grad,
                         int* iterations, double* residual)
{
    adjoint::CsrMatrix A;
    A.n = size_t(n);
    A.rowPtr.assign(rp, rp + n + 1);
    A.colIdx.assign(ci, ci + rp[n]);
    A.values.assign(v, v + rp[n]);
    adjoint::DiscreteAdjointInputs in;
    in.A = &A;
    in.dJdQ.assign(dJdQ, dJdQ + n);
    if (dJdAlpha)
        in.dJdAlpha.assign(dJdAlpha, dJdAlpha + m);
    // G psi needs psi: the reference calls the callable after solving, and
    // the result's psi is only returned afterwards, so the callable reads it
    // through this pointer once the solve has filled it.
    const std::vector<double>* psiRef = nullptr;
    if (G)
        in.dRdAlphaTransposeTimes = [&](std::vector<double>& out) {
            for (int i = 0; i < m; ++i) {
                double s = 0.0;
                for (int j = 0; j < n; ++j)
                    s += G[size_t(i) * size_t(n) + size_t(j)] * (psiRef ? (*psiRef)[size_t(j)] : 0.0);
                out[size_t(i)] = s;
            }
        };
    adjoint::DiscreteAdjointOptions opt;
    opt.maxIter = size_t(maxIter);
    opt.restart = size_t(restart);
    opt.tol = tol;
    // First pass solves for psi; the gradient callable then sees psi.
    auto r1 = adjoint::solve_discrete_adjoint(in, size_t(m), opt);
    psiRef = &r1.psi;
    auto r = adjoint::solve_discrete_adjoint(in, size_t(m), opt);
    std::memcpy(psi, r.psi.data(), size_t(n) * sizeof(double));
    if (grad)
        std::memcpy(grad, r.gradient.data(), size_t(m) * sizeof(double));
    *iterations = int(r.iterations);
    *residual = r.residual;
    return r.ok ? 0 : 1;
}

// CG (the loop of CGSolver::solve, src/solver/LinearSolvers.cpp:420-452)
// preconditioned by the reference's Ruge-Stueben AMGPreconditioner.
int ref_rs_amg_pcg(int n, const int32_t* rp, const int32_t* ci, const double* v, const double* b,
                   double tol, int maxIter, double* x, double* last, double* setup_seconds)
{
    solver::CSRMatrix A;
    A.rowPtr.assign(rp, rp + n + 1);
    A.colIdx.assign(ci, ci + rp[n]);
    A.values.assign(v, v + rp[n]);
    solver::AMGPreconditioner M;
    const auto t0 = std::chrono::steady_clock::now();
    M.setup(A);
    *setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    using Vec = util::aligned_vector<double>;
    const size_t nn = static_cast<size_t>(n);
    Vec X(nn, 0.0), r(b, b + n), z(nn), p, Ap(nn);
    auto dot = [&](const Vec& a, const Vec& c) {
        double s = 0;
        for (int i = 0; i < n; ++i)
            s += a[size_t(i)] * c[size_t(i)];
        return s;
    };
    M.apply(r, z);
    p = z;
    double rz = dot(r, z);
    Vec B(b, b + n);
    const double bnorm = std::max(std::sqrt(dot(B, B)), 1e-30);
    int it = 0;
    for (; it < maxIter; ++it) {
        A.spmv(p, Ap);
        const double pAp = dot(p, Ap);
        if (pAp == 0)
            break;
        const double alpha = rz / pAp;
        for (int i = 0; i < n; ++i) {
            X[size_t(i)] += alpha * p[size_t(i)];
            r[size_t(i)] -= alpha * Ap[size_t(i)];
        }
        *last = std::sqrt(dot(r, r)) / bnorm;
        if (*last < tol) {
            ++it;
            break;
        }
        M.apply(r, z);
        const double rzNew = dot(r, z);
        const double beta = rzNew / rz;
        for (int i = 0; i < n; ++i)
            p[size_t(i)] = z[size_t(i)] + beta * p[size_t(i)];
        rz = rzNew;
    }
    std::memcpy(x, X.data(), size_t(n) * sizeof(double));
    return it;
}

// LeastSquaresGradient::evaluate of phi on the reference mesh.
int ref_ls_gradient(int Nx, int Ny, int Nz, const double* o, const double* L, double perturb,
                    uint64_t seed, const double* phi, double* gx, double* gy, double* gz)
{
    auto m = make_mesh(Nx, Ny, Nz, o, L, perturb, seed);
    const size_t nC = m->cells().size();
    solver::ScalarField f(phi, phi + nC);
    solver::VectorField g;
    solver::LeastSquaresGradient G(*m);
    G.evaluate(f, g);
    std::memcpy(gx, g.x.data(), nC * sizeof(double));
    std::memcpy(gy, g.y.data(), nC * sizeof(double));
    std::memcpy(gz, g.z.data(), nC * sizeof(double));
    return 0;
}

// SimpleAlgorithm::correct_fields with the given aP, p', p, U and phi;
// p, U (3 n) and phi are updated in place.
int ref_correct_fields(int Nx, int Ny, int Nz, const double* o, const double* L, double perturb,
                       uint64_t seed, const int* bc_of_zone, int nz, double rho, double urfP,
                       const double* aP, const double* pp, double* p, double* U, double* phi)
{
    auto c = make_case(Nx, Ny, Nz, o, L, perturb, seed, bc_of_zone, nz, nullptr, rho, 1e-3, 0.0,
                       0.7, U);
    const size_t nC = c->mesh->cells().size(), nF = c->mesh->faces().size();
    auto* P = c->fields.find_scalar("p");
    std::memcpy(P->data(), p, nC * sizeof(double));
    c->probe->options().urfP = urfP;
    c->probe->correct(aP, pp, phi);
    auto* V = c->fields.find_vector("U");
    std::memcpy(p, P->data(), nC * sizeof(double));
    std::memcpy(U, V->x.data(), nC * sizeof(double));
    std::memcpy(U + nC, V->y.data(), nC * sizeof(double));
    std::memcpy(U + 2 * nC, V->z.data(), nC * sizeof(double));
    std::memcpy(phi, c->probe->phi().data(), nF * sizeof(double));
    return 0;
}

// n_steps calls of advance_time_step(n_inner) (n_steps = 0: one iterate())
// of SimpleAlgorithm (algorithm 0; 2: with the SIMPLEC denominators, as
// SimplecAlgorithm) or PisoAlgorithm (1) with a dense direct
// solve.  U (3 n) and p are the initial fields and receive the final ones;
// phi the final face flux; res[4] the last SimpleResiduals.
int ref_flow(int Nx, int Ny, int Nz, const double* o, const double* L, double perturb,
             uint64_t seed, const int* bc_of_zone, int nz, const double* bc_value, int algorithm,
             double rho, double mu, double urfU, double urfP, double dt, int bdf2,
             int n_correctors, int n_nonorth, int n_steps, int n_inner, double* U, double* p,
             double* phi, double* res)
{
    auto mesh = make_mesh(Nx, Ny, Nz, o, L, perturb, seed);
    const size_t nC = mesh->cells().size(), nF = mesh->faces().size();
    std::vector<solver::BoundarySpec> bcs;
    for (int z = 1; z < nz; ++z) {
        solver::BoundarySpec b;
        b.zone = meshing::ZoneId(z);
        b.type = bc_type(bc_of_zone[z]);
        for (int d = 0; d < 3; ++d)
            b.vectorValue[d] = bc_value ? bc_value[z * 3 + d] : 0.0;
        bcs.push_back(b);
    }
    solver::FieldRegistry fields;
    auto& Uf = fields.vector("U", nC);
    auto& pf = fields.scalar("p", nC);
    std::memcpy(Uf.x.data(), U, nC * sizeof(double));
    std::memcpy(Uf.y.data(), U + nC, nC * sizeof(double));
    std::memcpy(Uf.z.data(), U + 2 * nC, nC * sizeof(double));
    std::memcpy(pf.data(), p, nC * sizeof(double));
    DenseSolver sm, sp;
    const auto scheme = bdf2 ? solver::TemporalScheme::BDF2 : solver::TemporalScheme::ImplicitEuler;
    std::unique_ptr<solver::SimpleAlgorithm> alg;
    const util::aligned_vector<double>* phiRef = nullptr;
    if (algorithm == 1) {
        solver::PisoOptions po;
        po.nCorrectors = n_correctors;
        po.nNonOrthCorr = n_nonorth;
        po.rho = rho;
        po.mu = mu;
        po.dt = dt;
        po.timeScheme = scheme;
        auto a = std::make_unique<PisoProbe>(*mesh, fields, bcs, sm, sp, po);
        phiRef = &a->phi();
        alg = std::move(a);
    } else {
        solver::SimpleOptions so;
        so.rho = rho;
        so.mu = mu;
        so.urfU = urfU;
        so.urfP = urfP;
        so.dt = dt;
        so.timeScheme = scheme;
        if (algorithm == 2)
            so.algorithm = solver::PvCouplingVariant::SIMPLEC;
        auto a = std::make_unique<SimpleProbe>(*mesh, fields, bcs, sm, sp, so);
        phiRef = &a->phi();
        alg = std::move(a);
    }
    solver::SimpleResiduals r{};
    if (n_steps == 0)
        r = alg->iterate();
    for (int s = 0; s < n_steps; ++s)
        r = alg->advance_time_step(n_inner);
    auto& Uo = *fields.find_vector("U");
    std::memcpy(U, Uo.x.data(), nC * sizeof(double));
    std::memcpy(U + nC, Uo.y.data(), nC * sizeof(double));
    std::memcpy(U + 2 * nC, Uo.z.data(), nC * sizeof(double));
    std::memcpy(p, fields.find_scalar("p")->data(), nC * sizeof(double));
    std::memcpy(phi, phiRef->data(), nF * sizeof(double));
    for (int k = 0; k < 3; ++k)
        res[k] = r.mom[k];
    res[3] = r.cont;
    return 0;
}

// io::Checkpoint::write of the reference mesh with scalar field "p" and
// vector field "U" (n_cells values each; U as x, y, z blocks).
int ref_checkpoint_write(const char* path, int Nx, int Ny, int Nz, const double* o,
                         const double* L, double perturb, uint64_t seed, const double* p,
                         const double* U, int iteration, double time, uint64_t hash)
{
    auto m = make_mesh(Nx, Ny, Nz, o, L, perturb, seed);
    const size_t nC = m->cells().size();
    solver::FieldRegistry F;
    std::memcpy(F.scalar("p", nC).data(), p, nC * sizeof(double));
    auto& V = F.vector("U", nC);
    std::memcpy(V.x.data(), U, nC * sizeof(double));
    std::memcpy(V.y.data(), U + nC, nC * sizeof(double));
    std::memcpy(V.z.data(), U + 2 * nC, nC * sizeof(double));
    io::CheckpointMeta meta;
    meta.iteration = iteration;
    meta.time = time;
    meta.settingsHash = hash;
    return io::Checkpoint::write(path, *m, F, meta) ? 0 : 1;
}

// io::Checkpoint::read: 0 on success, 1 when the reader refuses the file.
// sizes[3] = cells, faces, nodes of the mesh read back; the mesh arrays go to
// the same buffers as ref_mesh_arrays; scalar "p" and vector "U" (3 n blocks)
// are copied when present (have[0], have[1]).
int ref_checkpoint_read(const char* path, int* iteration, double* time, uint64_t* hash,
                        int64_t* sizes, uint64_t* owner, uint64_t* neighbor, double* area,
                        double* fc, uint32_t* zone, double* vol, double* cc, int32_t* cfo,
                        uint64_t* cfi, double* p, double* U, int* have)
{
    meshing::Mesh m;
    solver::FieldRegistry F;
    io::CheckpointMeta meta;
    if (!io::Checkpoint::read(path, m, F, meta))
        return 1;
    *iteration = meta.iteration;
    *time = meta.time;
    *hash = meta.settingsHash;
    const auto& Fc = m.faces();
    const auto& C = m.cells();
    const size_t nF = Fc.owner.size(), nC = C.volume.size();
    sizes[0] = int64_t(nC);
    sizes[1] = int64_t(nF);
    sizes[2] = int64_t(m.nodes().x.size());
    if (owner) {
        std::memcpy(owner, Fc.owner.data(), nF * 8);
        std::memcpy(neighbor, Fc.neighbor.data(), nF * 8);
        std::memcpy(area, Fc.areaX.data(), nF * 8);
        std::memcpy(area + nF, Fc.areaY.data(), nF * 8);
        std::memcpy(area + 2 * nF, Fc.areaZ.data(), nF * 8);
        std::memcpy(fc, Fc.centroidX.data(), nF * 8);
        std::memcpy(fc + nF, Fc.centroidY.data(), nF * 8);
        std::memcpy(fc + 2 * nF, Fc.centroidZ.data(), nF * 8);
        std::memcpy(zone, Fc.boundaryZone.data(), nF * 4);
        std::memcpy(vol, C.volume.data(), nC * 8);
        std::memcpy(cc, C.centroidX.data(), nC * 8);
        std::memcpy(cc + nC, C.centroidY.data(), nC * 8);
        std::memcpy(cc + 2 * nC, C.centroidZ.data(), nC * 8);
        std::memcpy(cfo, C.faceOffsets.data(), (nC + 1) * 4);
        std::memcpy(cfi, C.faceIndices.data(), size_t(C.faceOffsets[nC]) * 8);
    }
    have[0] = have[1] = 0;
    if (auto* P = F.find_scalar("p")) {
        have[0] = 1;
        if (p)
            std::memcpy(p, P->data(), P->size() * sizeof(double));
    }
    if (auto* V = F.find_vector("U")) {
        have[1] = 1;
        if (U) {
            const size_t n = V->x.size();
            std::memcpy(U, V->x.data(), n * sizeof(double));
            std::memcpy(U + n, V->y.data(), n * sizeof(double));
            std::memcpy(U + 2 * n, V->z.data(), n * sizeof(double));
        }
    }
    return 0;
}

} // extern "C"
