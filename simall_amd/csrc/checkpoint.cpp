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

#include "mi_common.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

struct mi_checkpoint
{
    mi_checkpoint_meta meta{};
    uint32_t version = 0;
    int64_t n_nodes = 0, n_faces = 0, n_cells = 0;
    std::map<std::string, std::vector<double>> scalars;
    std::map<std::string, std::vector<double>> vectors[3];
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
        if (ok && n && std::fseek(f, long(n * elem), SEEK_CUR) != 0)
            ok = false;
        return n;
    }
    std::vector<double> doubles()
    {
        const uint64_t n = val<uint64_t>();
        std::vector<double> v;
        if (!ok || n > (uint64_t(1) << 40)) {
            ok = false;
            return v;
        }
        v.resize(size_t(n));
        bytes(v.data(), size_t(n) * sizeof(double));
        return v;
    }
    std::string str()
    {
        const uint32_t n = val<uint32_t>();
        std::string s;
        if (!ok || n > (1u << 20)) {
            ok = false;
            return s;
        }
        s.resize(n);
        bytes(s.data(), n);
        return s;
    }
};

constexpr char kMagic[8] = {'S', 'I', 'M', 'A', 'L', 'L', 'C', 'P'};

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
    auto c = new (std::nothrow) mi_checkpoint();
    if (!c) {
        std::fclose(in.f);
        return set_error(MI_ERR_DEVICE, "out of host memory");
    }
    char magic[8];
    in.bytes(magic, 8);
    c->version = in.val<uint32_t>();
    (void)in.val<uint32_t>(); // flags
    (void)in.val<uint64_t>(); // timestamp
    int rc = MI_OK;
    if (!in.ok || std::memcmp(magic, kMagic, 8) != 0) {
        rc = set_error(MI_ERR_BAD_ARG, "%s: not a SIMALLCP checkpoint", path);
    } else if (c->version > 1) {
        rc = set_error(MI_ERR_UNSUPPORTED, "%s: checkpoint version %u not supported", path,
                       c->version);
    } else {
        c->meta.iteration = in.val<int32_t>();
        (void)in.val<uint32_t>();
        c->meta.time = in.val<double>();
        c->meta.settings_hash = in.val<uint64_t>();
        // mesh section: counts only
        c->n_nodes = int64_t(in.skip(8));
        in.skip(8);
        in.skip(8);
        c->n_faces = int64_t(in.skip(8)); // owner
        in.skip(8);                        // neighbour
        for (int k = 0; k < 6; ++k)
            in.skip(8); // areas, centroids
        in.skip(4);     // boundary zone
        in.skip(4);     // node offsets
        in.skip(8);     // node indices
        c->n_cells = int64_t(in.skip(8)); // volume
        for (int k = 0; k < 3; ++k)
            in.skip(8);
        in.skip(4);
        in.skip(8);
        const uint32_t ns = in.val<uint32_t>(), nv = in.val<uint32_t>();
        for (uint32_t i = 0; in.ok && i < ns; ++i) {
            std::string name = in.str();
            c->scalars[name] = in.doubles();
        }
        for (uint32_t i = 0; in.ok && i < nv; ++i) {
            std::string name = in.str();
            for (int d = 0; d < 3; ++d)
                c->vectors[d][name] = in.doubles();
        }
        if (!in.ok)
            rc = set_error(MI_ERR_BAD_ARG, "%s: truncated or malformed checkpoint", path);
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

MI_API int mi_checkpoint_close(mi_checkpoint* c)
{
    delete c;
    return MI_OK;
}

} // extern "C"
