// mesh.hpp — the SoA mesh on the device (mi_mesh) and the view kernels take.
#pragma once

#include "csr.hpp"

#include <memory>
#include <vector>

namespace mi
{

enum : int8_t { BC_WALL = 0, BC_PRESSURE = 1, BC_SYMMETRY = 2, BC_VELOCITY_INLET = 3, BC_OUTFLOW = 4,
                BC_PERIODIC = 5 };

// sl_tag bits
constexpr int kSlotOwner = 16;    // the slot's cell owns the face
constexpr int kSlotPeriodic = 32; // a boundary face with a periodic twin (sl_oth = twin's cell)

struct MeshDev
{
    int n_cells, n_faces;
    const int* owner;
    const int* nbr; // -1 on boundary faces
    const double *ax, *ay, *az, *fcx, *fcy, *fcz;
    const int8_t* fbc;  // boundary class of each face (interior: -1)
    const int* ftwin;   // the owner of a boundary face's periodic twin, else -1
    const int* fzone;   // zone id of each face
    const double *ccx, *ccy, *ccz, *vol;
    const int* cf_off;  // [n_cells+1]
    const int* cf_face; // faces of each cell, ascending id
    const int* cf_slot; // CSR position of (cell, other cell); -1 on boundary faces
    const int* rp;      // CSR row pointer (for zeroing rows)
    const int* diag;    // CSR position of the diagonal
    // per cell-face slot (the cf_* order): |S_f|^2, |d . S_f| (d from the
    // owner's centroid to the neighbour's, or to the face centroid on a
    // boundary), the cell across the face (-1 on the boundary; the twin's
    // cell on a periodic face), and a tag: kSlotOwner when the slot's cell
    // owns the face, kSlotPeriodic on a periodic face, plus the face's
    // boundary class + 1
    const double* sl_s2;
    const double* sl_ds;
    const int* sl_oth;
    const int8_t* sl_tag;
    // The same slot arrays sliced for the cell-per-thread assembly kernels:
    // slot j of cell c at ss_off[c / 64] + 64 j + c % 64 (each slice of 64
    // cells padded to its longest cell), so slot j of a wavefront's cells is
    // one coalesced access; ss_cnt[c] = the cell's slot count, negated when
    // its slots do not write distinct off-diagonal entries of its row (the
    // assembly kernels then zero the row and accumulate into it).
    const int* ss_off;
    const int* ss_cnt;
    const int* ss_face;
    const int* ss_slot;
    const int* ss_oth;
    const int8_t* ss_tag;
    const double* ss_s2;
    const double* ss_ds;
};

constexpr int kSlotSlice = 64;

} // namespace mi

struct mi_mesh
{
    std::shared_ptr<mi::Pattern> pat;
    int64_t n_cells = 0, n_faces = 0;
    int n_zones = 0;
    std::vector<int> bc_of_zone;
    bool any_dirichlet = false;
    mi::DevBuf<int> owner, nbr, fzone, cf_off, cf_face, cf_slot, sl_oth, ftwin;
    std::vector<int> twin_face; // periodic twin face of every face, -1 (host)
    mi::DevBuf<int8_t> fbc, sl_tag;
    mi::DevBuf<double> sl_s2, sl_ds;
    // sliced copies (MeshDev::ss_*)
    mi::DevBuf<int> ss_off, ss_cnt, ss_face, ss_slot, ss_oth;
    mi::DevBuf<int8_t> ss_tag;
    mi::DevBuf<double> ss_s2, ss_ds;
    mi::DevBuf<double> ax, ay, az, fcx, fcy, fcz, ccx, ccy, ccz, vol;
    // LeastSquaresGradient's inverted 3x3 moment matrices, 6 per cell, built
    // on first use (geometry only)
    mi::DevBuf<double> ls_minv;
    // Face-gradient state, built on first use: cells renumbered in the order
    // the face list first meets them (fg_perm[new] = old), every face's owner
    // and neighbour in that numbering (neighbour -1: pressure boundary, -2:
    // a boundary that contributes nothing, -3 - t: a periodic face whose
    // twin's cell is t), |S_f|^2 and |d_f . S_f|; and the
    // renumbered copies of x, lambda and aP a call gathers.
    mi::DevBuf<int> fg_perm, fg_o, fg_n;
    mi::DevBuf<double> fg_s2, fg_ds, fg_x, fg_l, fg_a;
    int fg_pin = -1; // cell 0 in the renumbering
    mi::MeshDev dev() const
    {
        mi::MeshDev d{};
        d.n_cells = int(n_cells);
        d.n_faces = int(n_faces);
        d.owner = owner.p;
        d.nbr = nbr.p;
        d.ax = ax.p;
        d.ay = ay.p;
        d.az = az.p;
        d.fcx = fcx.p;
        d.fcy = fcy.p;
        d.fcz = fcz.p;
        d.fbc = fbc.p;
        d.ftwin = ftwin.p;
        d.fzone = fzone.p;
        d.ccx = ccx.p;
        d.ccy = ccy.p;
        d.ccz = ccz.p;
        d.vol = vol.p;
        d.cf_off = cf_off.p;
        d.cf_face = cf_face.p;
        d.cf_slot = cf_slot.p;
        d.rp = pat->rp.p;
        d.diag = pat->diag.p;
        d.sl_s2 = sl_s2.p;
        d.sl_ds = sl_ds.p;
        d.sl_oth = sl_oth.p;
        d.sl_tag = sl_tag.p;
        d.ss_off = ss_off.p;
        d.ss_cnt = ss_cnt.p;
        d.ss_face = ss_face.p;
        d.ss_slot = ss_slot.p;
        d.ss_oth = ss_oth.p;
        d.ss_tag = ss_tag.p;
        d.ss_s2 = ss_s2.p;
        d.ss_ds = ss_ds.p;
        return d;
    }
};

namespace mi
{
// Asynchronous launchers of the assembly kernels on stream s (assembly.hip).
int launch_simplec_denominators(const mi_mesh* M, const double* amom, const double* aP,
                                double* den, hipStream_t s);
int launch_assemble_pressure(const mi_mesh* M, double rho, const double* aP, const double* gamma,
                             const double* phi, int pin, double* val, double* rhs, hipStream_t s);
int launch_assemble_momentum(const mi_mesh* M, int comp, double mu, double rho, const double* phi,
                             const double* pgradV, const double* dt, const double* Uold,
                             const double* Unm1, const double* Ucomp, double urf,
                             const double* bc_dev, double* val, double* rhs, hipStream_t s);
int launch_face_flux(const mi_mesh* M, double rho, const double* Ux, const double* Uy,
                     const double* Uz, const double* bc_dev, double* phi, hipStream_t s);
// Zone velocity table [n_zones * 3] on the device (h nullable: zeros).
int mesh_bc_values(const mi_mesh* M, const double* h, DevBuf<double>& d, hipStream_t s);
} // namespace mi
