/*
 * workload.h — synthetic inputs for the pressure-correction solve path.
 *
 * Builds the SoA hex meshes and per-face / per-cell data of the five solver
 * workloads (128x128x1 cavity slab, perturbed and permuted N^3 pressure
 * Poisson, Taylor-Green momentum, batched transport, adjoint).  The mesh
 * mirrors the reference's Cartesian generator and connectivity builder:
 *   - i-fastest node and cell numbering (src/meshing/CartesianMesher.cpp:21-60)
 *   - faces emitted per cell as -Z,+Z,-Y,+Y,-X,+X, de-duplicated so that the
 *     first cell seen owns the face (src/meshing/Connectivity.cpp:74-90)
 *   - node-average face centroid, fan-summed area vector, divergence-theorem
 *     volume and centroid (src/meshing/MeshStorage.cpp:21-97)
 *   - boundary zones 1..6 = xMin,xMax,yMin,yMax,zMin,zMax
 *     (src/meshing/CartesianMesher.cpp:64-89).
 * Interior nodes may be perturbed (before geometry) and cells renumbered by
 * a seeded permutation (after geometry); face order never changes.
 *
 * Random numbers: SplitMix64, uniform = (x >> 11) * 2^-53, normal by the
 * cosine branch of Box-Muller, Fisher-Yates permutation.
 */
#ifndef SIMALL_WORKLOAD_H
#define SIMALL_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t n_cells, n_faces, n_nodes;
    /* nodes */
    double *nodeX, *nodeY, *nodeZ;
    /* faces (meshing::FaceStorage layout; neighbor UINT64_MAX = boundary) */
    uint64_t *owner, *neighbor;
    double *areaX, *areaY, *areaZ;
    double *faceCx, *faceCy, *faceCz;
    uint32_t* boundaryZone;
    int32_t* faceNodeOffsets;  /* n_faces + 1 */
    uint64_t* faceNodeIndices; /* 4 per face */
    /* cells */
    double* volume;
    double *cellCx, *cellCy, *cellCz;
    int32_t* cellFaceOffsets;  /* n_cells + 1 */
    uint64_t* cellFaceIndices; /* 6 per cell */
} wl_mesh;

/* Generates an Nx*Ny*Nz hex mesh of the box [o, o+L].  perturb > 0 moves
 * every interior node by U[-perturb*h, perturb*h] per coordinate (h = the
 * spacing along that axis) from seed node_seed.  perm_seed != 0 renumbers
 * cells by a Fisher-Yates permutation.  Returns 0 on success. */
int wl_mesh_cartesian(int Nx, int Ny, int Nz, double ox, double oy, double oz,
                      double Lx, double Ly, double Lz, double perturb,
                      uint64_t node_seed, uint64_t perm_seed, wl_mesh* out);
void wl_mesh_free(wl_mesh* m);

/* Fills out[0..n) from SplitMix64 seeded with `seed`. */
void wl_fill_uniform(uint64_t seed, double lo, double hi, double* out, int64_t n);
void wl_fill_normal(uint64_t seed, double* out, int64_t n);
/* out_i = 10^U[lo_exp, hi_exp] */
void wl_fill_loguniform(uint64_t seed, double lo_exp, double hi_exp, double* out, int64_t n);
/* Subtracts the arithmetic mean (left-to-right sum). */
void wl_remove_mean(double* v, int64_t n);

/* Taylor-Green velocity u = sin x cos y cos z, v = -cos x sin y cos z, w = 0
 * at cell centroids (applications/simall_taylor_green/main.cpp:114-115). */
void wl_taylor_green(const wl_mesh* m, double* Ux, double* Uy, double* Uz);

/* Random permutation of [0,n) (Fisher-Yates, SplitMix64). */
void wl_permutation(uint64_t seed, int64_t n, int64_t* perm);

#ifdef __cplusplus
}
#endif

#endif
