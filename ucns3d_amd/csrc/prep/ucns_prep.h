/*
 * ucns_prep.h -- host-side FP64 preprocessor (libucnsprep.so).
 *
 * Builds everything the residual path reads but never writes: the mesh
 * (synthetic boxes, periodic or bounded per axis), face connectivity, geometry, stencils and the
 * PRESTORE arrays (pseudo-inverses, smoothness-indicator matrices, basis
 * integrals, face Gauss points in each cell's local frame).  It restates the
 * reference's setup chain:
 *   NEIGHBOURSS / GEOMETRY_CALC   src/grid_p.f90:1850, src/grid_t.f90:1192-1560
 *   DETSTEN + ALLSf (BFS)         src/grid_p.f90:3979-4551
 *   STENCIILS + CHECK_CONDITION   src/grid_p.f90:5822-6090, 4582-4811
 *   STENCILS (ADMIS, FULL)        src/grid_p.f90:6452-6560
 *   LOCALISE_STENCIL / _STEN2     src/local_pt.f90:653-1230
 *   CHECKGRADS (face reorient)    src/local_pt.f90:2220-2286
 *   FIND_ROT_ANGLES               src/grid_t.f90:274-456
 *   PRESTORE_RECONSTRUCTION3      src/prestore.f90:167-560, 1529-2025, 3625-4069
 *   QRDECOMPOSITION               src/matrix.f90:93-146
 *   MEMORY_FAST (face qps)        src/reconstruct_p.f90:45-302
 * and fills the ucns_*_v1 structs of include/ucns.h, so the GPU library
 * receives exactly what a Fortran host would pass.
 */
#ifndef UCNS_PREP_H
#define UCNS_PREP_H

#include <stdint.h>
#include "../../../include/ucns.h"

#ifdef __cplusplus
extern "C" {
#endif

/* element shapes (IELEM%ISHAPE) */
#define UCNSP_HEX    1
#define UCNSP_TET    2
#define UCNSP_PRISM  3
#define UCNSP_PYRAMID 4
#define UCNSP_HYBRID 5
#define UCNSP_CHANNEL 6
#define UCNSP_BLAST  7

typedef struct ucnsp_opts {
    int32_t shape;          /* UCNSP_HEX: nx*ny*nz hexahedra; UCNSP_TET: 6 Kuhn tets per hex;
                               UCNSP_PRISM: 2 prisms per hex; UCNSP_PYRAMID: 6 pyramids
                               per hex (its faces as bases, its centre the apex);
                               UCNSP_HYBRID: hexes, pyramids, prisms in thirds of x;
                               UCNSP_CHANNEL: 2 prisms per hex in the wall layers
                               j < min(10, ny/3), then one layer of 5 pyramids + 2 tets
                               per hex, then 6 Kuhn tets per hex;
                               UCNSP_BLAST: a central box of hexahedra (the middle half of
                               each axis), a one-hex shell around it of pyramids (faces
                               towards hexahedra and shell hexes) and tetrahedron pairs
                               (faces towards tetrahedra and prisms), 2 prisms per hex in
                               the layers k < max(1, nz/10) and k >= nz - max(1, nz/10),
                               6 Kuhn tets per hex elsewhere */
    int32_t nx, ny, nz;
    double  lx, ly, lz;     /* periodic box [0,lx]x[0,ly]x[0,lz] */
    double  jitter;         /* interior vertices moved by U(-j,j)*spacing per coordinate */
    uint64_t seed;
    int32_t iorder;         /* 1 MUSCL (linear), 3 WENO-4 (cubic) ... */
    int32_t iweno;          /* 1 WENO, -1 MUSCL */
    int32_t extf;           /* stencil extension factor (profile 8: 2) */
    int32_t weight_lsqr;    /* 0 unweighted, 1 inverse-distance */
    int32_t nthreads;       /* OpenMP threads, <=0: library default */
    int32_t bc[6];          /* boundary code of the x-, x+, y-, y+, z-, z+ sides; 0 on both sides of an
                               axis makes it periodic (code 5 on the paired faces).  Codes as IBOUND%ICODE:
                               1 inflow, 2 outflow, 3 symmetry, 4 wall, 6 farfield */
    double  bump[8];        /* lower-wall bump a_k: with x^ = 4x/lx, z^ = z/lz, the nodes move by
                               dy = ly (1 - y/ly) sum_k a_k sin(k pi (x^-1)/2) sin(pi z^) for 1 <= x^ <= 3 */
    int32_t ees;            /* 0 WENO; 5 CWENO: also the low-order sector data (ucnsp_fill_cweno) */
} ucnsp_opts;

typedef struct ucnsp_case ucnsp_case;

int  ucnsp_build(const ucnsp_opts *opts, ucnsp_case **out);
/* GRID.cel / GRID.vrt / GRID.bnd (OPEN_ARBITRARY and the readers of
 * src/io.f90:4157-4301, src/grid_p.f90:1655-1760, src/bc_p.f90:30-140): text
 * (binary = 0) or Fortran unformatted sequential records (binary = 1, BINIO);
 * cells as 8 node ids whose repeats give the shape, nodes as id x y z,
 * boundary faces as id + 4 node ids (a triangle repeats its third) + code.
 * ucnsp_write_grid writes a case's mesh (periodic pairs on both sides, code 5);
 * ucnsp_read_grid builds a case from the files with the scheme options of
 * opts (its shape, box, jitter, bc and bump are ignored): periodic axes from
 * the faces of codes 5/50, paired with their translates across the bounding
 * box. */
int  ucnsp_write_grid(const ucnsp_case *c, const char *dir, int32_t binary);
int  ucnsp_read_grid(const char *dir, int32_t binary, const ucnsp_opts *opts, ucnsp_case **out);
/* RESTART.dat of an explicit march (CHECKPOINT, src/io.f90:13457-13639): step
 * (int32), time (float64), ncell x 5 cell averages (float64) in global cell order */
int  ucnsp_write_restart(const char *path, int32_t it, double t, const double *u, int64_t ncell);
int  ucnsp_read_restart(const char *path, int64_t ncell, int32_t *it, double *t, double *u);
/* Fill the boundary structs with pointers into the case's storage (valid
 * until ucnsp_free).  precision/gamma/lwci1 are the caller's scheme choices. */
int  ucnsp_fill(const ucnsp_case *c, int32_t precision, double gamma, double lwci1,
                ucns_mesh_v2 *mesh, ucns_scheme_v2 *scheme, ucns_prestore_v1 *pre);
/* The CWENO arrays of a case built with ees 5 (pointers into its storage). */
int  ucnsp_fill_cweno(const ucnsp_case *c, ucns_cweno_v1 *cw);
/* Named internal arrays for tests and drivers: "xyz", "xyz_box" (before
 * the bump), "cell_nodes", "centre", "ihexl", "anglex", "angley", ... ;
 * dtype 0=int32, 1=float64. */
int  ucnsp_array(const ucnsp_case *c, const char *name, void **ptr, int64_t *count, int32_t *dtype);
/* Integer / real metadata: ints[0..15], dbls[0..7] (see ucns_prep.cpp). */
int  ucnsp_info(const ucnsp_case *c, int64_t *ints, double *dbls);
/* Cell averages (AoS [ncell][5], conserved) of the periodic vortex
 * rho=1(+noise), u=sin x cos y cos z, v=-cos x sin y cos z, w=0,
 * p=100/gamma+(rho/16)(cos2x+cos2y)(cos2z+2).  noise_amp*N(0,1) is added to
 * each cell's density (seeded). */
int  ucnsp_init_vortex(const ucnsp_case *c, double gamma, double noise_amp, uint64_t seed, double *u_aos);
/* Element tables of one cell shape (UCNSP_HEX/TET/PRISM/PYRAMID) into
 * out[64]: [0] nodes, [1] faces, [2+f] nodes of face f, [8+4f+k] its local
 * node k, [32] DECOMPOSE3 sub-tetrahedra, [33+4t+k] their local nodes; -1
 * where unused. */
int  ucnsp_shape_table(int32_t shape, int32_t *out);
/* Gauss-point tables used by the face integration (weights sum to 1). */
int  ucnsp_face_rule(int32_t igqrules, int32_t tri, double *pts /*[n][3] barycentric or (r,s,0)*/, double *w, int32_t *n);
/* Move the nodes to xyz [nnode][3] and rebuild the geometry and PRESTORE
 * data for them with the topology, stencils and FULL flags unchanged (a
 * mesh deformation, as a shape study or its finite differences need: the
 * stencil selection's cone tests are not continuous in the nodes).  Data are
 * then solved cell by cell, never copied between congruent cells. */
int  ucnsp_move_nodes(ucnsp_case *c, const double *xyz);
void ucnsp_free(ucnsp_case *c);
const char *ucnsp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
