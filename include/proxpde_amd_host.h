/*
 * proxpde_amd host mesh / dof-map layer (C ABI), the host-side mirror of the
 * reference's setup path:
 *   buildHyperCube / buildLine / buildSquare / buildCube  src/mesh.cpp:17-535
 *   buildFacets + markFacetsCube                          include/mesh.hpp:226-481
 *   DOF::setupElemMap (first-visit numbering)             include/dof.hpp:112-329
 *   BCEss::setupDofMap (dofs of marked facets)            include/bc.hpp:253-295
 *   dof points used by interpolateAnalyticFunction        include/fespace.hpp:195-238
 *   readGMSH (GMSH 2.2 ASCII)                             include/mesh.hpp:785-1062
 * Pure host code, no device memory.  Status codes as in proxpde_amd.h.
 */
#ifndef PROXPDE_AMD_HOST_H
#define PROXPDE_AMD_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* structured mesh kinds */
#define PXH_MESH_LINE 1      /* buildLine                              */
#define PXH_MESH_SQUARE 2    /* buildSquare<Quad>                      */
#define PXH_MESH_CUBE_HEX 3  /* buildCube<Hexahedron>                  */
#define PXH_MESH_CUBE_TET5 4 /* buildCube<Tetrahedron>: 5 tets/cube    */
#define PXH_MESH_CUBE_TET6 5 /* Kuhn split, 6 tets/cube (no reference counterpart) */

/* boundary markers, reference include/mesh.hpp:178-188 */
#define PXH_SIDE_BOTTOM 1
#define PXH_SIDE_RIGHT 2
#define PXH_SIDE_TOP 3
#define PXH_SIDE_LEFT 4
#define PXH_SIDE_BACK 5
#define PXH_SIDE_FRONT 6

typedef struct pxh_mesh pxh_mesh;

int pxh_mesh_structured(pxh_mesh ** mesh, int kind, const double * origin, const double * length, int nx, int ny,
                        int nz);

/* element geometries of an unstructured mesh */
#define PXH_GEO_LINE 1
#define PXH_GEO_QUAD 2
#define PXH_GEO_HEX 3
#define PXH_GEO_TET 4
/* GMSH 2.2 ASCII mesh file, readGMSH (include/mesh.hpp:785-1062): the volume
 * elements of geometry `geo` (quadratic ones read through their vertices),
 * the points they use renumbered in file order, every boundary facet marked
 * with the first tag of the facet element of the file on the same vertices
 * (0 when the file lists none).  Errors: PXG_E_INVALID + pxh_last_error.  */
int pxh_mesh_gmsh(pxh_mesh ** mesh, const char * path, int geo);
/* message of the last failed host-layer call                              */
const char * pxh_last_error(void);
int pxh_mesh_destroy(pxh_mesh * mesh);
/* sizes: points, elements, points per element, boundary facets */
int pxh_mesh_info(const pxh_mesh * mesh, int64_t * n_pts, int64_t * n_elem, int * npe, int64_t * n_bfacets);
int pxh_mesh_get(const pxh_mesh * mesh, double * coords, uint32_t * conn);
int pxh_mesh_set_coords(pxh_mesh * mesh, const double * coords);
/* boundary facets as (element, side, marker) triples */
int pxh_mesh_bfacets(const pxh_mesh * mesh, int32_t * out);

/* first-visit dof numbering for fe_type (PXG_FE_*): call with elem_map NULL
 * to get n_dofs and the local dof count, then with an n_elem x n buffer */
int pxh_dofmap(const pxh_mesh * mesh, int fe_type, int64_t * n_dofs, int * n_local, uint32_t * elem_map);
/* sorted, unique dofs on facets marked `marker`; *count in/out (capacity / size) */
int pxh_bc_dofs(const pxh_mesh * mesh, int fe_type, const uint32_t * elem_map, int marker, uint32_t * dofs,
                int64_t * count);
/* dof coordinates (n_dofs x 3) */
int pxh_dof_coords(const pxh_mesh * mesh, int fe_type, const uint32_t * elem_map, int64_t n_dofs, double * out);

#ifdef __cplusplus
}
#endif

#endif /* PROXPDE_AMD_HOST_H */
