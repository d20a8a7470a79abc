// checkpoint.cpp — the SIMALLCP checkpoint file of the reference
// (Checkpoint::write / Checkpoint::read, src/io/Checkpoint.cpp:67-230;
// layout declared in src/io/Checkpoint.hpp:27-42), host side.
//
// File: header {char magic[8] = "SIMALLCP"; u32 version = 1; u32 flags;
// u64 timestamp}, meta {i32 iteration; 4 bytes of padding; f64 time;
// u64 settings hash}, then the mesh section (nodes, faces, cells) and the
// field section (u32 scalar count, u32 vector count, then per field a
// u32-length name and u64-length arrays), little endian.  Every array is a
// u64 element count followed by the elements.  Written to <path>.tmp and
// renamed over <path>, as the reference does.
#include "simall_amd.h"

#include "checkpoint.hpp"
#include "mi_common.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

struct mi_checkpoint
{
    mi_checkpoint_meta meta{};
    uint32_t version = 0;
    int64_t n_nodes = 0, n_faces = 0, n_cells = 0;
    std::map<std::string, std::vector<double>> scalars;
    std::map<std::string, std::vector<double>> vectors[3];
    std::unique_ptr<mi::AmgHierarchyHost> amg; // optional trailer
};

namespace
{

using mi::set_error;

struct Out
{
    FILE* f = nullptr;
    bool ok = true;
    void bytes(const void* p, size_t n)
    {
        if (ok && n && std::fwrite(p, 1, n, f) != n)
            ok = false;
    }
    template <class T> void val(const T& v) { bytes(&v, sizeof(T)); }
    template <class T> void vec(const T* p, uint64_t n)
    {
        val(n);
        bytes(p, size_t(n) * sizeof(T));
    }
    void str(const char* s)
    {
        const uint32_t n = uint32_t(std::strlen(s));
        val(n);
        bytes(s, n);
    }
};

struct In
{
    FILE* f = nullptr;
    bool ok = true;
    uint64_t size = 0; // file size: no array may claim more than what is left
    bool fits(uint64_t n, size_t elem)
    {
        const long pos = std::ftell(f);
        if (pos < 0 || n > (uint64_t(1) << 40) || n * elem > size - uint64_t(pos))
            ok = false;
        return ok;
    }
    void bytes(void* p, size_t n)
    {
        if (ok && n && std::fread(p, 1, n, f) != n)
            ok = false;
    }
    template <class T> T val()
    {
        T v{};
        bytes(&v, sizeof(T));
        return v;
    }
    // reads a u64 count and skips the elements; returns the count
    uint64_t skip(size_t elem)
    {
        const uint64_t n = val<uint64_t>();
        if (ok)
            fits(n, elem);
        if (ok && n && std::fseek(f, long(n * elem), SEEK_CUR) != 0)
            ok = false;
        return n;
    }
    std::vector<double> doubles()
    {
        const uint64_t n = val<uint64_t>();
        std::vector<double> v;
        if (!ok || !fits(n, sizeof(double)))
            return v;
        v.resize(size_t(n));
        bytes(v.data(), size_t(n) * sizeof(double));
        return v;
    }
    std::string str()
    {
        const uint32_t n = val<uint32_t>();
        std::string s;
        if (!ok || n > (1u << 20) || !fits(n, 1))
            return s;
        s.resize(n);
        bytes(s.data(), n);
        return s;
    }
};

constexpr char kMagic[8] = {'S', 'I', 'M', 'A', 'L', 'L', 'C', 'P'};
constexpr char kAmgTag[8] = {'S', 'I', 'M', 'A', 'M', 'G', 'H', '1'};

// AMG trailer: tag, u32 levels, per level {u32 has_P, CSR A, f64 lambda_max,
// i32 aggregates, [CSR P, CSR R]}, i32 coarse n, dense coarsest operator
// (u64 count + f64), f64 operator and grid complexity.  A CSR is i32 rows,
// i32 columns, then rowPtr, colIdx, values as u64 count + elements.
void put_csr(Out& o, const mi::HCsr& a)
{
    o.val(int32_t(a.nr));
    o.val(int32_t(a.nc));
    o.vec(a.rp.data(), uint64_t(a.rp.size()));
    o.vec(a.ci.data(), uint64_t(a.ci.size()));
    o.vec(a.v.data(), uint64_t(a.v.size()));
}

template <class T> bool get_vec(In& in, std::vector<T>& v, uint64_t expect)
{
    const uint64_t n = in.val<uint64_t>();
    if (!in.ok || n != expect || !in.fits(n, sizeof(T))) {
        in.ok = false;
        return false;
    }
    v.resize(size_t(n));
    in.bytes(v.data(), size_t(n) * sizeof(T));
    return in.ok;
}

bool get_csr(In& in, mi::HCsr& a)
{
    a.nr = in.val<int32_t>();
    a.nc = in.val<int32_t>();
    if (!in.ok || a.nr < 0 || a.nc < 0 || !get_vec(in, a.rp, uint64_t(a.nr) + 1))
        return false;
    const int nnz = a.rp.back();
    if (a.rp.front() != 0 || nnz < 0 || !get_vec(in, a.ci, uint64_t(nnz)) ||
        !get_vec(in, a.v, uint64_t(nnz)))
        return false;
    for (int i = 0; i < a.nr; ++i)
        if (a.rp[size_t(i) + 1] < a.rp[size_t(i)])
            return false;
    for (int c : a.ci)
        if (c < 0 || c >= a.nc)
            return false;
    return true;
}

bool get_amg(In& in, mi::AmgHierarchyHost& h)
{
    const uint32_t nl = in.val<uint32_t>();
    if (!in.ok || nl == 0 || nl > 64)
        return false;
    h.levels.assign(nl, mi::AmgLevelHost{});
    for (uint32_t l = 0; l < nl; ++l) {
        mi::AmgLevelHost& L = h.levels[l];
        const uint32_t has_p = in.val<uint32_t>();
        if (!get_csr(in, L.A))
            return false;
        L.lmax = in.val<double>();
        L.n_aggregates = in.val<int32_t>();
        if (has_p && (!get_csr(in, L.P) || !get_csr(in, L.R)))
            return false;
        if (!in.ok || (has_p != 0) != (l + 1 < nl))
            return false;
    }
    // The V-cycle indexes level l+1's vectors with R_l's rows and P_l's
    // columns: a trailer whose shapes do not chain is refused before any of
    // it reaches the device.
    for (uint32_t l = 0; l < nl; ++l) {
        const mi::AmgLevelHost& L = h.levels[l];
        if (L.A.nr != L.A.nc)
            return false;
        if (l + 1 < nl) {
            const int nf = L.A.nr, nc = h.levels[l + 1].A.nr;
            if (L.P.nr != nf || L.P.nc != nc || L.R.nr != nc || L.R.nc != nf)
                return false;
        }
    }
    h.coarse_n = in.val<int32_t>();
    if (!in.ok || h.coarse_n != h.levels.back().A.nr ||
        !get_vec(in, h.coarse_dense, uint64_t(h.coarse_n) * uint64_t(h.coarse_n)))
        return false;
    h.operator_complexity = in.val<double>();
    h.grid_complexity = in.val<double>();
    return in.ok;
}

} // namespace

namespace mi
{

int checkpoint_append_amg(const char* path, const AmgHierarchyHost& h)
{
    if (!path || h.levels.empty() ||
        h.coarse_dense.size() != size_t(h.coarse_n) * size_t(h.coarse_n))
        return set_error(MI_ERR_BAD_ARG, "no hierarchy to append");
    Out o;
    o.f = std::fopen(path, "ab");
    if (!o.f)
        return set_error(MI_ERR_BAD_ARG, "cannot open %s", path);
    o.bytes(kAmgTag, 8);
    o.val(uint32_t(h.levels.size()));
    for (size_t l = 0; l < h.levels.size(); ++l) {
        const AmgLevelHost& L = h.levels[l];
        const bool has_p = l + 1 < h.levels.size();
        o.val(uint32_t(has_p ? 1 : 0));
        put_csr(o, L.A);
        o.val(L.lmax);
        o.val(int32_t(L.n_aggregates));
        if (has_p) {
            put_csr(o, L.P);
            put_csr(o, L.R);
        }
    }
    o.val(int32_t(h.coarse_n));
    o.vec(h.coarse_dense.data(), uint64_t(h.coarse_dense.size()));
    o.val(h.operator_complexity);
    o.val(h.grid_complexity);
    const bool wrote = o.ok;
    if (std::fclose(o.f) != 0 || !wrote)
        return set_error(MI_ERR_DEVICE, "append to %s failed", path);
    return MI_OK;
}

const AmgHierarchyHost* checkpoint_amg(const mi_checkpoint* c)
{
    return c ? c->amg.get() : nullptr;
}

} // namespace mi

namespace
{

// Header, meta, the mesh section (counts only), the fields and the optional
// AMG trailer of an open checkpoint file.
int read_checkpoint(In& in, mi_checkpoint& c, const char* path)
{
    char magic[8];
    in.bytes(magic, 8);
    c.version = in.val<uint32_t>();
    (void)in.val<uint32_t>(); // flags
    (void)in.val<uint64_t>(); // timestamp
    if (!in.ok || std::memcmp(magic, kMagic, 8) != 0)
        return set_error(MI_ERR_BAD_ARG, "%s: not a SIMALLCP checkpoint", path);
    if (c.version > 1)
        return set_error(MI_ERR_UNSUPPORTED, "%s: checkpoint version %u not supported", path,
                         c.version);
    c.meta.iteration = in.val<int32_t>();
    (void)in.val<uint32_t>();
    c.meta.time = in.val<double>();
    c.meta.settings_hash = in.val<uint64_t>();
    // mesh section: counts only
    c.n_nodes = int64_t(in.skip(8));
    in.skip(8);
    in.skip(8);
    c.n_faces = int64_t(in.skip(8)); // owner
    in.skip(8);                       // neighbour
    for (int k = 0; k < 6; ++k)
        in.skip(8); // areas, centroids
    in.skip(4);     // boundary zone
    in.skip(4);     // node offsets
    in.skip(8);     // node indices
    c.n_cells = int64_t(in.skip(8)); // volume
    for (int k = 0; k < 3; ++k)
        in.skip(8);
    in.skip(4);
    in.skip(8);
    const uint32_t ns = in.val<uint32_t>(), nv = in.val<uint32_t>();
    for (uint32_t i = 0; in.ok && i < ns; ++i) {
        std::string name = in.str();
        c.scalars[name] = in.doubles();
    }
    for (uint32_t i = 0; in.ok && i < nv; ++i) {
        std::string name = in.str();
        for (int d = 0; d < 3; ++d)
            c.vectors[d][name] = in.doubles();
    }
    if (!in.ok)
        return set_error(MI_ERR_BAD_ARG, "%s: truncated or malformed checkpoint", path);
    // optional AMG trailer
    char tag[8];
    if (std::fread(tag, 1, 8, in.f) != 8)
        return MI_OK;
    if (std::memcmp(tag, kAmgTag, 8) != 0)
        return set_error(MI_ERR_BAD_ARG, "%s: unknown data after the field section", path);
    c.amg = std::make_unique<mi::AmgHierarchyHost>();
    if (!get_amg(in, *c.amg))
        return set_error(MI_ERR_BAD_ARG, "%s: truncated or malformed AMG trailer", path);
    char extra;
    if (std::fread(&extra, 1, 1, in.f) == 1)
        return set_error(MI_ERR_BAD_ARG, "%s: unknown data after the AMG trailer", path);
    return MI_OK;
}

} // namespace

extern "C" {

MI_API int mi_checkpoint_write(const char* path, const mi_mesh_soa* m, const mi_mesh_nodes* nodes,
                               int n_scalars, const mi_named_field* scalars, int n_vectors,
                               const mi_named_vector* vectors, const mi_checkpoint_meta* meta)
{
    mi::clear_error();
    if (!path || !m || !nodes || n_scalars < 0 || n_vectors < 0 ||
        (n_scalars && !scalars) || (n_vectors && !vectors))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    const uint64_t nN = uint64_t(nodes->n_nodes), nF = uint64_t(m->n_faces),
                   nC = uint64_t(m->n_cells);
    for (int i = 0; i < n_scalars; ++i)
        if (!scalars[i].name || (nC && !scalars[i].data))
            return set_error(MI_ERR_BAD_ARG, "scalar field %d: null name or data", i);
    for (int i = 0; i < n_vectors; ++i)
        if (!vectors[i].name || (nC && (!vectors[i].x || !vectors[i].y || !vectors[i].z)))
            return set_error(MI_ERR_BAD_ARG, "vector field %d: null name or data", i);
    const std::string tmp = std::string(path) + ".tmp";
    Out o;
    o.f = std::fopen(tmp.c_str(), "wb");
    if (!o.f)
        return set_error(MI_ERR_BAD_ARG, "cannot open %s", tmp.c_str());
    // header and meta (the reference writes both structs as they lie in memory)
    o.bytes(kMagic, 8);
    o.val(uint32_t(1));
    o.val(uint32_t(0));
    o.val(uint64_t(std::chrono::duration_cast<std::chrono::seconds>(
                       std::chrono::system_clock::now().time_since_epoch())
                       .count()));
    const mi_checkpoint_meta mz = meta ? *meta : mi_checkpoint_meta{};
    o.val(int32_t(mz.iteration));
    o.val(uint32_t(0));
    o.val(mz.time);
    o.val(uint64_t(mz.settings_hash));
    // mesh
    o.vec(nodes->x, nN);
    o.vec(nodes->y, nN);
    o.vec(nodes->z, nN);
    o.vec(m->owner, nF);
    o.vec(m->neighbor, nF);
    o.vec(m->areaX, nF);
    o.vec(m->areaY, nF);
    o.vec(m->areaZ, nF);
    o.vec(m->faceCx, nF);
    o.vec(m->faceCy, nF);
    o.vec(m->faceCz, nF);
    o.vec(m->boundaryZone, nF);
    o.vec(nodes->faceNodeOffsets, nF + 1);
    o.vec(nodes->faceNodeIndices, nF ? uint64_t(nodes->faceNodeOffsets[nF]) : 0);
    o.vec(m->volume, nC);
    o.vec(m->cellCx, nC);
    o.vec(m->cellCy, nC);
    o.vec(m->cellCz, nC);
    o.vec(m->cellFaceOffsets, nC + 1);
    o.vec(m->cellFaceIndices, nC ? uint64_t(m->cellFaceOffsets[nC]) : 0);
    // fields
    o.val(uint32_t(n_scalars));
    o.val(uint32_t(n_vectors));
    for (int i = 0; i < n_scalars; ++i) {
        o.str(scalars[i].name);
        o.vec(scalars[i].data, nC);
    }
    for (int i = 0; i < n_vectors; ++i) {
        o.str(vectors[i].name);
        o.vec(vectors[i].x, nC);
        o.vec(vectors[i].y, nC);
        o.vec(vectors[i].z, nC);
    }
    const bool wrote = o.ok;
    if (std::fclose(o.f) != 0 || !wrote) {
        std::remove(tmp.c_str());
        return set_error(MI_ERR_DEVICE, "write to %s failed", tmp.c_str());
    }
    if (std::rename(tmp.c_str(), path) != 0)
        return set_error(MI_ERR_DEVICE, "rename %s -> %s failed", tmp.c_str(), path);
    return MI_OK;
}

MI_API int mi_checkpoint_open(const char* path, mi_checkpoint** out)
{
    mi::clear_error();
    if (!path || !out)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    *out = nullptr;
    In in;
    in.f = std::fopen(path, "rb");
    if (!in.f)
        return set_error(MI_ERR_BAD_ARG, "cannot open %s", path);
    if (std::fseek(in.f, 0, SEEK_END) == 0) {
        const long end = std::ftell(in.f);
        in.size = end > 0 ? uint64_t(end) : 0;
    }
    std::rewind(in.f);
    auto c = new (std::nothrow) mi_checkpoint();
    int rc = MI_OK;
    if (!c) {
        rc = set_error(MI_ERR_DEVICE, "out of host memory");
    } else {
        try {
            rc = read_checkpoint(in, *c, path);
        } catch (const std::bad_alloc&) {
            // sizes read from a damaged file
            rc = set_error(MI_ERR_BAD_ARG, "%s: sizes in the checkpoint exceed host memory", path);
        }
    }
    std::fclose(in.f);
    if (rc != MI_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return MI_OK;
}

MI_API int mi_checkpoint_info(const mi_checkpoint* c, mi_checkpoint_meta* meta, int64_t* n_cells,
                              int64_t* n_faces, int64_t* n_nodes)
{
    if (!c)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    if (meta)
        *meta = c->meta;
    if (n_cells)
        *n_cells = c->n_cells;
    if (n_faces)
        *n_faces = c->n_faces;
    if (n_nodes)
        *n_nodes = c->n_nodes;
    return MI_OK;
}

MI_API int mi_checkpoint_scalar(const mi_checkpoint* c, const char* name, double* out,
                                int64_t capacity, int64_t* count)
{
    if (!c || !name)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    const auto it = c->scalars.find(name);
    if (it == c->scalars.end())
        return set_error(MI_ERR_BAD_ARG, "no scalar field '%s' in the checkpoint", name);
    if (count)
        *count = int64_t(it->second.size());
    if (out)
        std::memcpy(out, it->second.data(),
                    sizeof(double) * size_t(std::min<int64_t>(capacity, int64_t(it->second.size()))));
    return MI_OK;
}

MI_API int mi_checkpoint_vector(const mi_checkpoint* c, const char* name, double* x, double* y,
                                double* z, int64_t capacity, int64_t* count)
{
    if (!c || !name)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    const auto it = c->vectors[0].find(name);
    if (it == c->vectors[0].end())
        return set_error(MI_ERR_BAD_ARG, "no vector field '%s' in the checkpoint", name);
    double* dst[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
        const std::vector<double>& v = c->vectors[d].at(name);
        if (d == 0 && count)
            *count = int64_t(v.size());
        if (dst[d])
            std::memcpy(dst[d], v.data(),
                        sizeof(double) * size_t(std::min<int64_t>(capacity, int64_t(v.size()))));
    }
    return MI_OK;
}

MI_API int mi_checkpoint_amg_levels(const mi_checkpoint* c, int* levels)
{
    if (!c || !levels)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    *levels = c->amg ? int(c->amg->levels.size()) : 0;
    return MI_OK;
}

MI_API int mi_checkpoint_append_amg(const char* path, const mi_amg_host* h)
{
    mi::clear_error();
    const mi::AmgHierarchyHost* H = mi::amg_host_hierarchy(h);
    if (!H)
        return set_error(MI_ERR_BAD_ARG, "null hierarchy");
    return mi::checkpoint_append_amg(path, *H);
}

MI_API int mi_checkpoint_close(mi_checkpoint* c)
{
    delete c;
    return MI_OK;
}

} // extern "C"
