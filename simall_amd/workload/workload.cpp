// workload.cpp — synthetic meshes and fields for the solver workloads.
// Compiled with -ffp-contract=off so geometry rounds like the reference's
// serial x86-64 code (no fused multiply-add).
#include "workload.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace
{

uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double uniform01(uint64_t& s)
{
    return double(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

double normal01(uint64_t& s)
{
    const double u1 = 1.0 - uniform01(s);
    const double u2 = uniform01(s);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
}

template <class T> T* alloc(int64_t n)
{
    return static_cast<T*>(std::calloc(size_t(n > 0 ? n : 1), sizeof(T)));
}

struct V3
{
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

} // namespace

extern "C" void wl_mesh_free(wl_mesh* m)
{
    if (!m)
        return;
    void* ptrs[] = {m->nodeX, m->nodeY, m->nodeZ, m->owner, m->neighbor, m->areaX, m->areaY,
                    m->areaZ, m->faceCx, m->faceCy, m->faceCz, m->boundaryZone, m->faceNodeOffsets,
                    m->faceNodeIndices, m->volume, m->cellCx, m->cellCy, m->cellCz,
                    m->cellFaceOffsets, m->cellFaceIndices};
    for (void* p : ptrs)
        std::free(p);
    std::memset(m, 0, sizeof(*m));
}

extern "C" int wl_mesh_cartesian(int Nx, int Ny, int Nz, double ox, double oy, double oz,
                                 double Lx, double Ly, double Lz, double perturb,
                                 uint64_t node_seed, uint64_t perm_seed, wl_mesh* out)
{
    if (!out || Nx < 1 || Ny < 1 || Nz < 1)
        return 1;
    std::memset(out, 0, sizeof(*out));
    const int64_t Px = Nx + 1, Py = Ny + 1, Pz = Nz + 1;
    const int64_t nN = Px * Py * Pz;
    const int64_t nC = int64_t(Nx) * Ny * Nz;
    const int64_t nF = 3 * nC + int64_t(Nx) * Ny + int64_t(Ny) * Nz + int64_t(Nx) * Nz;
    if (nC >= (int64_t(1) << 31))
        return 2;
    wl_mesh& m = *out;
    m.n_nodes = nN;
    m.n_cells = nC;
    m.n_faces = nF;

    // ---- nodes (src/meshing/CartesianMesher.cpp:16-30)
    const double dx = Lx / Nx, dy = Ly / Ny, dz = Lz / Nz;
    m.nodeX = alloc<double>(nN);
    m.nodeY = alloc<double>(nN);
    m.nodeZ = alloc<double>(nN);
    for (int64_t k = 0; k < Pz; ++k)
        for (int64_t j = 0; j < Py; ++j)
            for (int64_t i = 0; i < Px; ++i) {
                const int64_t id = (k * Py + j) * Px + i;
                m.nodeX[id] = ox + int(i) * dx;
                m.nodeY[id] = oy + int(j) * dy;
                m.nodeZ[id] = oz + int(k) * dz;
            }
    if (perturb > 0.0) {
        uint64_t s = node_seed;
        for (int64_t k = 0; k < Pz; ++k)
            for (int64_t j = 0; j < Py; ++j)
                for (int64_t i = 0; i < Px; ++i) {
                    const bool interior = i > 0 && i < Nx && j > 0 && j < Ny && k > 0 && k < Nz;
                    if (!interior)
                        continue;
                    const int64_t id = (k * Py + j) * Px + i;
                    m.nodeX[id] += (2.0 * uniform01(s) - 1.0) * perturb * dx;
                    m.nodeY[id] += (2.0 * uniform01(s) - 1.0) * perturb * dy;
                    m.nodeZ[id] += (2.0 * uniform01(s) - 1.0) * perturb * dz;
                }
    }

    // ---- faces and cell->face CSR (src/meshing/Connectivity.cpp:62-101).
    // A face is new unless the adjacent cell across it has a smaller index
    // (already visited in i-fastest order); that cell then owns it.
    m.owner = alloc<uint64_t>(nF);
    m.neighbor = alloc<uint64_t>(nF);
    m.boundaryZone = alloc<uint32_t>(nF);
    m.faceNodeOffsets = alloc<int32_t>(nF + 1);
    m.faceNodeIndices = alloc<uint64_t>(4 * nF);
    m.cellFaceOffsets = alloc<int32_t>(nC + 1);
    m.cellFaceIndices = alloc<uint64_t>(6 * nC);
    std::vector<int64_t> fPlusX(static_cast<size_t>(nC)), fPlusY(static_cast<size_t>(nC)),
        fPlusZ(static_cast<size_t>(nC));
    auto pid = [&](int64_t i, int64_t j, int64_t k) { return uint64_t((k * Py + j) * Px + i); };
    auto cid = [&](int64_t i, int64_t j, int64_t k) { return (k * Ny + j) * Nx + i; };
    int64_t nf = 0;
    auto newFace = [&](uint64_t a, uint64_t b, uint64_t c, uint64_t d, int64_t cell) {
        const int64_t f = nf++;
        m.faceNodeOffsets[f] = int32_t(4 * f);
        m.faceNodeIndices[4 * f + 0] = a;
        m.faceNodeIndices[4 * f + 1] = b;
        m.faceNodeIndices[4 * f + 2] = c;
        m.faceNodeIndices[4 * f + 3] = d;
        m.owner[f] = uint64_t(cell);
        m.neighbor[f] = UINT64_MAX;
        m.boundaryZone[f] = 1;
        return f;
    };
    for (int64_t k = 0; k < Nz; ++k)
        for (int64_t j = 0; j < Ny; ++j)
            for (int64_t i = 0; i < Nx; ++i) {
                const int64_t c = cid(i, j, k);
                const uint64_t n0 = pid(i, j, k), n1 = pid(i + 1, j, k), n2 = pid(i + 1, j + 1, k),
                               n3 = pid(i, j + 1, k), n4 = pid(i, j, k + 1),
                               n5 = pid(i + 1, j, k + 1), n6 = pid(i + 1, j + 1, k + 1),
                               n7 = pid(i, j + 1, k + 1);
                uint64_t* cf = m.cellFaceIndices + 6 * c;
                m.cellFaceOffsets[c] = int32_t(6 * c);
                int64_t f;
                // -Z {n0,n3,n2,n1}
                if (k == 0)
                    f = newFace(n0, n3, n2, n1, c);
                else {
                    f = fPlusZ[size_t(cid(i, j, k - 1))];
                    m.neighbor[f] = uint64_t(c);
                }
                cf[0] = uint64_t(f);
                // +Z {n4,n5,n6,n7}
                f = newFace(n4, n5, n6, n7, c);
                fPlusZ[size_t(c)] = f;
                cf[1] = uint64_t(f);
                // -Y {n0,n1,n5,n4}
                if (j == 0)
                    f = newFace(n0, n1, n5, n4, c);
                else {
                    f = fPlusY[size_t(cid(i, j - 1, k))];
                    m.neighbor[f] = uint64_t(c);
                }
                cf[2] = uint64_t(f);
                // +Y {n3,n7,n6,n2}
                f = newFace(n3, n7, n6, n2, c);
                fPlusY[size_t(c)] = f;
                cf[3] = uint64_t(f);
                // -X {n0,n4,n7,n3}
                if (i == 0)
                    f = newFace(n0, n4, n7, n3, c);
                else {
                    f = fPlusX[size_t(cid(i - 1, j, k))];
                    m.neighbor[f] = uint64_t(c);
                }
                cf[4] = uint64_t(f);
                // +X {n1,n2,n6,n5}
                f = newFace(n1, n2, n6, n5, c);
                fPlusX[size_t(c)] = f;
                cf[5] = uint64_t(f);
            }
    m.faceNodeOffsets[nf] = int32_t(4 * nf);
    m.cellFaceOffsets[nC] = int32_t(6 * nC);
    if (nf != nF)
        return 3;
    for (int64_t f = 0; f < nF; ++f)
        m.boundaryZone[f] = (m.neighbor[f] == UINT64_MAX) ? 1u : 0u;

    // ---- geometry (src/meshing/MeshStorage.cpp:21-97)
    m.areaX = alloc<double>(nF);
    m.areaY = alloc<double>(nF);
    m.areaZ = alloc<double>(nF);
    m.faceCx = alloc<double>(nF);
    m.faceCy = alloc<double>(nF);
    m.faceCz = alloc<double>(nF);
    for (int64_t f = 0; f < nF; ++f) {
        const int32_t s = m.faceNodeOffsets[f], e = m.faceNodeOffsets[f + 1];
        V3 c{0, 0, 0};
        for (int32_t i = s; i < e; ++i) {
            const uint64_t n = m.faceNodeIndices[i];
            c.x += m.nodeX[n];
            c.y += m.nodeY[n];
            c.z += m.nodeZ[n];
        }
        const double inv = 1.0 / double(e - s);
        c.x *= inv;
        c.y *= inv;
        c.z *= inv;
        m.faceCx[f] = c.x;
        m.faceCy[f] = c.y;
        m.faceCz[f] = c.z;
        V3 area{0, 0, 0};
        for (int32_t i = s; i < e; ++i) {
            const uint64_t a = m.faceNodeIndices[i];
            const uint64_t b = m.faceNodeIndices[(i + 1 < e) ? i + 1 : s];
            const V3 pa{m.nodeX[a], m.nodeY[a], m.nodeZ[a]};
            const V3 pb{m.nodeX[b], m.nodeY[b], m.nodeZ[b]};
            area = add(area, mul(cross(sub(pa, c), sub(pb, c)), 0.5));
        }
        m.areaX[f] = area.x;
        m.areaY[f] = area.y;
        m.areaZ[f] = area.z;
    }
    m.volume = alloc<double>(nC);
    m.cellCx = alloc<double>(nC);
    m.cellCy = alloc<double>(nC);
    m.cellCz = alloc<double>(nC);
    for (int64_t c = 0; c < nC; ++c) {
        double V = 0;
        V3 Cw{0, 0, 0};
        for (int32_t i = m.cellFaceOffsets[c]; i < m.cellFaceOffsets[c + 1]; ++i) {
            const uint64_t f = m.cellFaceIndices[i];
            const V3 xf{m.faceCx[f], m.faceCy[f], m.faceCz[f]};
            const V3 af{m.areaX[f], m.areaY[f], m.areaZ[f]};
            const double sgn = (m.owner[f] == uint64_t(c)) ? 1.0 : -1.0;
            const double dv = sgn * dot(xf, af) / 3.0;
            V += dv;
            Cw = add(Cw, mul(mul(xf, dv), 3.0 / 4.0));
        }
        m.volume[c] = V;
        if (V != 0) {
            m.cellCx[c] = Cw.x / V;
            m.cellCy[c] = Cw.y / V;
            m.cellCz[c] = Cw.z / V;
        }
    }

    // ---- boundary zones (src/meshing/CartesianMesher.cpp:64-89)
    const double tol = 1e-9;
    for (int64_t f = 0; f < nF; ++f) {
        if (m.neighbor[f] != UINT64_MAX)
            continue;
        const double cx = m.faceCx[f], cy = m.faceCy[f], cz = m.faceCz[f];
        if (std::abs(cx - ox) < tol)
            m.boundaryZone[f] = 1;
        else if (std::abs(cx - (ox + Lx)) < tol)
            m.boundaryZone[f] = 2;
        else if (std::abs(cy - oy) < tol)
            m.boundaryZone[f] = 3;
        else if (std::abs(cy - (oy + Ly)) < tol)
            m.boundaryZone[f] = 4;
        else if (std::abs(cz - oz) < tol)
            m.boundaryZone[f] = 5;
        else if (std::abs(cz - (oz + Lz)) < tol)
            m.boundaryZone[f] = 6;
    }

    // ---- optional cell renumbering: new id = perm[old]
    if (perm_seed != 0) {
        std::vector<int64_t> perm(static_cast<size_t>(nC));
        wl_permutation(perm_seed, nC, perm.data());
        for (int64_t f = 0; f < nF; ++f) {
            m.owner[f] = uint64_t(perm[size_t(m.owner[f])]);
            if (m.neighbor[f] != UINT64_MAX)
                m.neighbor[f] = uint64_t(perm[size_t(m.neighbor[f])]);
        }
        auto permute = [&](double*& a) {
            double* b = alloc<double>(nC);
            for (int64_t c = 0; c < nC; ++c)
                b[perm[size_t(c)]] = a[c];
            std::free(a);
            a = b;
        };
        permute(m.volume);
        permute(m.cellCx);
        permute(m.cellCy);
        permute(m.cellCz);
        uint64_t* cfi = alloc<uint64_t>(6 * nC);
        for (int64_t c = 0; c < nC; ++c)
            std::memcpy(cfi + 6 * perm[size_t(c)], m.cellFaceIndices + 6 * c, 6 * sizeof(uint64_t));
        std::free(m.cellFaceIndices);
        m.cellFaceIndices = cfi;
    }
    return 0;
}

extern "C" void wl_permutation(uint64_t seed, int64_t n, int64_t* perm)
{
    for (int64_t i = 0; i < n; ++i)
        perm[i] = i;
    uint64_t s = seed;
    for (int64_t i = n - 1; i > 0; --i) {
        const int64_t j = int64_t(splitmix64(s) % uint64_t(i + 1));
        const int64_t t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
    }
}

extern "C" void wl_fill_uniform(uint64_t seed, double lo, double hi, double* out, int64_t n)
{
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i)
        out[i] = lo + (hi - lo) * uniform01(s);
}

extern "C" void wl_fill_normal(uint64_t seed, double* out, int64_t n)
{
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i)
        out[i] = normal01(s);
}

extern "C" void wl_fill_loguniform(uint64_t seed, double lo_exp, double hi_exp, double* out,
                                   int64_t n)
{
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i)
        out[i] = std::pow(10.0, lo_exp + (hi_exp - lo_exp) * uniform01(s));
}

extern "C" void wl_remove_mean(double* v, int64_t n)
{
    if (n <= 0)
        return;
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i)
        s += v[i];
    const double mean = s / double(n);
    for (int64_t i = 0; i < n; ++i)
        v[i] -= mean;
}

extern "C" void wl_taylor_green(const wl_mesh* m, double* Ux, double* Uy, double* Uz)
{
    for (int64_t c = 0; c < m->n_cells; ++c) {
        const double x = m->cellCx[c], y = m->cellCy[c], z = m->cellCz[c];
        Ux[c] = std::sin(x) * std::cos(y) * std::cos(z);
        Uy[c] = -std::cos(x) * std::sin(y) * std::cos(z);
        Uz[c] = 0.0;
    }
}
