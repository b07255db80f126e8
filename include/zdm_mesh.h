/*
 * zdm_mesh.h -- host-side finite-volume mesh service (C ABI).
 *
 * Replaces the DMPlex services the FV time step consumes:
 *   - cell/face connectivity and the cone -> support adjacency
 *     (reference VirtualSolver.hpp:386-511, DMPlexDistribute + DMSetBasicAdjacency at :439);
 *   - FV geometry (DMPlexGetGeometryFVM): cell centroid + volume, face centroid and the
 *     area-weighted normal oriented from support[0] to support[1] (outward on the boundary)
 *     (consumed at TransportStep.hpp:405-408, 434-435, 556-559; Gradient.hpp:92-93);
 *   - the "Face Sets" label (boundary tags) from Gmsh $PhysicalNames
 *     (MeshConfigLoader.hpp:15-41, VirtualSolver.hpp:394-436).
 *
 * Cells are optionally renumbered (reverse Cuthill-McKee or a Hilbert curve) so that face
 * neighbours are close in memory; faces are numbered by ascending (support[0], support[1]).
 * Every function returns 0 on success and a negative code on failure; the message is
 * available from zdm_mesh_last_error().
 */
#ifndef ZDM_MESH_H
#define ZDM_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZDM_MESH_MAX_FACES 4   /* segments: 2, triangles: 3, quadrilaterals: 4 */

typedef struct zdm_mesh {
    int dim;          /* 1 (segments) or 2 (triangles / quadrilaterals) */
    int ncells;
    int nfaces;
    int max_faces;    /* max faces per cell in this mesh */
    int nverts;
    int max_cell_verts;

    /* cells, in the (renumbered) solver order */
    double* cell_centroid;   /* ncells*3 */
    double* cell_volume;     /* ncells   */
    int*    cell_nfaces;     /* ncells   */
    int*    cell_faces;      /* ncells*max_faces, ascending face ids, -1 padded */
    int*    cell_nbrs;       /* ncells*max_faces, cell across that face, -1 on the boundary */
    int*    cell_perm;       /* ncells: solver cell -> input cell index */
    int*    cell_verts;      /* ncells*max_cell_verts, -1 padded (solver order) */

    /* faces, ordered by (support[0], support[1]) */
    int*    face_cells;      /* nfaces*2: support[0] < support[1]; support[1] = -1 on the boundary */
    double* face_normal;     /* nfaces*3: area-weighted normal, support[0] -> support[1] */
    double* face_centroid;   /* nfaces*3 */
    int*    face_tag;        /* nfaces: "Face Sets" value of boundary faces, -1 if untagged / interior */

    double* verts;           /* nverts*3 */

    /* Gmsh physical names: name i has tag physical_tag[i] and dimension physical_dim[i] */
    int   nphysical;
    int*  physical_tag;
    int*  physical_dim;
    char* physical_names;    /* nphysical entries of 64 chars, NUL terminated */
} zdm_mesh;

/* Build from an explicit cell list.  cell_verts is ncells*nvpc (-1 padded for mixed
 * meshes).  Boundary face tags: nbf faces given by their vertices (1 vertex per face in 1-D,
 * 2 in 2-D) with a tag each.  If mark_all_tag > 0, every boundary face without a tag gets
 * it (DMPlexMarkBoundaryFaces, VirtualSolver.hpp:428-436).  renumber != 0 applies RCM. */
int zdm_mesh_from_cells(int dim, int nverts, const double* xyz, int ncells, int nvpc,
                        const int* cell_verts, int nbf, const int* bface_verts,
                        const int* bface_tags, int mark_all_tag, int renumber,
                        zdm_mesh** out);

/* Unit square [0,1]^2 as an nx x ny grid of squares, each split into two right triangles
 * along the SW -> NE diagonal.  Interior nodes are jittered by U(-jitter, jitter) times the
 * edge length in each direction (SplitMix64 stream seeded with `seed`).  All boundary faces
 * get tag 1. */
int zdm_mesh_tri_grid(int nx, int ny, double jitter, uint64_t seed, int renumber,
                      zdm_mesh** out);

/* Uniform 1-D grid of n segments on [x0, x1]; left end tag 1, right end tag 2. */
int zdm_mesh_line(int n, double x0, double x1, int renumber, zdm_mesh** out);

/* Gmsh MSH 2.2 ASCII reader (points, lines, triangles, quadrilaterals). */
int zdm_mesh_read_msh(const char* path, int renumber, zdm_mesh** out);

/* Disjoint union of n meshes (an ensemble): member i's cells are [cell_offsets[i],
 * cell_offsets[i+1]) in the union, in the member's own order, and its faces follow the
 * previous members' faces.  cell_offsets (n+1 entries) may be NULL.  Physical names are
 * dropped (face tags are kept).  Stands in for the reference's one-DM-per-run setup when
 * many runs share one launch. */
int zdm_mesh_concat(int n, const zdm_mesh* const* meshes, int* cell_offsets, zdm_mesh** out);

void zdm_mesh_destroy(zdm_mesh* m);
const char* zdm_mesh_last_error(void);

/* Initial-condition files (IOManager::LoadSolution, IOManager.hpp:122-192, reads the "state"
 * Vec with PETSc's HDF5 viewer: (ncells, n_dof) float64 in the mesh file's cell order, or flat).
 * libhdf5 (>= 1.10) is opened at run time: $ZDM_HDF5_LIB alone when set, else libhdf5.so,
 * libhdf5.so.103, .200, .310, /opt/conda/lib/libhdf5.so; without it both calls fail with a
 * message naming what was tried.
 * zdm_h5_read: the dataset's rank (<= 4) and dims; with data non-NULL also its values,
 * converted to float64 (capacity = room in data, in values). */
int zdm_h5_read(const char* path, const char* dataset, double* data, uint64_t capacity, int* ndims,
                uint64_t* dims);
/* zdm_h5_write: a new file (an existing one is truncated) holding one float64 dataset. */
int zdm_h5_write(const char* path, const char* dataset, const double* data, int ndims, const uint64_t* dims);

/* SplitMix64: next value of the stream and a uniform double in [0,1). */
uint64_t zdm_splitmix64_next(uint64_t* state);
double   zdm_splitmix64_uniform(uint64_t* state);

#ifdef __cplusplus
}
#endif
#endif
