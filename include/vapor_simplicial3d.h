/* vapor_simplicial3d.h - C ABI of the simplicial 3D compute backend on gfx950.
 *
 * vapor's tetrahedral DEC solver (src/solvers/simplicial3d) hands three of its
 * stages to an optional backend, SimplicialCompute3D
 * (SimplicialCompute3D.hpp:7-13):
 *
 *   advectCirculation  trace every dual vertex back along the velocity and
 *                      integrate the traced dual edges around each primal
 *                      edge's dual face (SimplicialFluidSolver3D.cpp:426-450);
 *   advectDensity      trace every primal vertex back and sample the smoke
 *                      there, times the decay (SimplicialFluidSolver3D.cpp:595-614);
 *   recover            solve the edge-potential Poisson problem by
 *                      preconditioned CG (SimplicialFluidSolver3D.cpp:135-217).
 *
 * vapor's own GPU backend (gpu/GpuSimplicial3D.cpp) runs them as OpenGL
 * compute shaders. This library runs them as HIP kernels: vsx_backend_create
 * takes the mesh, its dual and the operator once, as flat arrays; each stage
 * call then takes the solver state that changes per step. The C++ binding
 * that fills these arrays from vapor's solver, in place of
 * gpu/GpuSimplicial3D.cpp, is vapor_gpu_simplicial3d.hpp.
 *
 * Errors follow vapor_mac3d.h: every call returns VGX_OK or a status code, the
 * message is vgx_last_error(). The reference's exception types map to
 * VGX_EINVAL (std::invalid_argument) and VGX_EDEVICE (std::runtime_error);
 * for a runtime_error the message is the reference's own text. */
#ifndef VAPOR_SIMPLICIAL3D_H
#define VAPOR_SIMPLICIAL3D_H

#include <stdint.h>

#include "vapor_mac3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Recovery order (vsx_backend_set_recovery).
 * VSX_RECOVER_REFERENCE: the CPU solver's CG step for step: its
 *   preconditioner (the Jacobi diagonal on boxes, the IC(0) factor where one
 *   is given) and its inner products summed in index order, so the result is
 *   the CPU's bit for bit.
 * VSX_RECOVER_FAST: Jacobi preconditioning and tree-ordered inner products;
 *   deterministic, within rounding of the reference order, far fewer
 *   synchronisations. This is the default. */
#define VSX_RECOVER_FAST 0
#define VSX_RECOVER_REFERENCE 1

/* The mesh, its circumcentric dual and the potential operator, as vapor's
 * simplicial3d::Mesh (SimplicialMesh3D.hpp:18-61), DualMesh
 * (SimplicialDual3D.hpp:22-45) and SimplicialFluidSolver3D
 * (SimplicialFluidSolver3D.hpp:86-104) hold them. Index arrays are int32,
 * points are three doubles (x, y, z). Offsets arrays are CSR: entries of row
 * r are [offsets[r], offsets[r+1]). The library copies everything during
 * vsx_backend_create; the caller's arrays may be freed afterwards. */
typedef struct vsx_mesh {
    int32_t n_vertices, n_tets, n_faces, n_edges;
    int32_t box_domain; /* Mesh::boxDomain */
    double size;        /* Mesh::size (also DualMesh::size_) */

    const double* vertices;      /* 3 per vertex */
    const int32_t* tet_vertices; /* 4 per tet (Tet::vertices) */
    const double* tet_gradients; /* 9 per tet: Tet::gradients[1..3] */

    /* Mesh::bins_ / boundaryBins_ over bins_per_axis^3 cells; bins_per_axis
     * 0 means no bins (Mesh::locate then scans every tet). */
    int32_t bins_per_axis;
    const int32_t* bin_offsets;  /* bins_per_axis^3 + 1 */
    const int32_t* bin_items;    /* tet ids */
    const int32_t* wall_offsets; /* bins_per_axis^3 + 1 */
    const int32_t* wall_items;   /* boundary face ids */

    const int32_t* face_vertices;     /* 3 per face */
    const int32_t* face_edges;        /* 3 per face */
    const int32_t* face_signs;        /* 3 per face, +1 or -1 */
    const int32_t* face_tets;         /* 2 per face, -1 on the boundary */
    const double* face_area_normals;  /* 3 per face (Face::areaNormal) */
    const int32_t* edge_vertices;     /* 2 per edge */
    const uint8_t* edge_boundary;     /* 1 per edge */
    const uint8_t* vertex_boundary;   /* 1 per vertex (Mesh::boundaryVertices) */
    const double* star0;              /* n_vertices */
    const double* star1;              /* n_edges */
    const double* star2;              /* n_faces */

    /* The solver's Jacobi diagonal (diagonal_) and, where it built one, its
     * IC(0) factor (incompleteLower_ as CSR over n_edges rows with ascending
     * columns, incompleteDiagonal_); ic_offsets NULL when there is none. */
    const double* diagonal;
    const int32_t* ic_offsets;
    const int32_t* ic_columns;
    const double* ic_values;
    const double* ic_diagonal;

    /* The dual (DualMesh::vertices, edges, loops). */
    int32_t n_dual, n_dual_edges;
    const double* dual_positions;  /* 3 per dual vertex */
    const int32_t* dual_edges;     /* 2 per dual edge (a, b) */
    const int32_t* loop_offsets;   /* n_edges + 1 */
    const int32_t* loop_edges;     /* dual edge ids */
    const int32_t* loop_signs;     /* +1 or -1 */

    /* Boxes: one convex Voronoi cell per primal vertex (DualMesh::cells). */
    const int32_t* cell_plane_offsets;  /* n_vertices + 1 */
    const double* cell_planes;          /* 4 per plane: normal, offset */
    const int32_t* cell_corner_offsets; /* n_vertices + 1 */
    const int32_t* corner_vertices;     /* dual vertex id per corner */
    const int32_t* corner_cone_offsets; /* n_corners + 1 */
    const int32_t* cone_planes;         /* 3 per cone, plane index within the cell */

    /* General domains: tet hints (dualTets_, primalTets_) and the
     * circumcentric subdivision (DualMesh::nodes_, subdivisions_). */
    int32_t n_nodes;
    const int32_t* dual_hints;     /* n_dual */
    const int32_t* primal_hints;   /* n_vertices */
    const double* node_positions;  /* 3 per node */
    const int32_t* sub_offsets;    /* n_tets + 1 */
    const int32_t* sub_nodes;      /* 4 per subtet */
    const double* sub_gradients;   /* 9 per subtet */
} vsx_mesh;

typedef struct vsx_backend vsx_backend;

/* GpuSimplicial3D::initialize (gpu/GpuSimplicial3D.cpp:54-108): checks every
 * index against its range and uploads the mesh. */
int vsx_backend_create(const vsx_mesh* mesh, vsx_backend** out);
void vsx_backend_destroy(vsx_backend* backend);
/* GpuSimplicial3D::device (gpu/GpuSimplicial3D.hpp:14): the GPU's name. */
const char* vsx_backend_device(const vsx_backend* backend);
int vsx_backend_set_recovery(vsx_backend* backend, int order);
int vsx_backend_recovery(const vsx_backend* backend);
/* The order backends created afterwards start with (process-wide). */
int vsx_set_default_recovery(int order);
int vsx_default_recovery(void);

/* SimplicialCompute3D::advectCirculation: dual_velocity holds the solver's
 * dualVelocity_ (3 per dual vertex), node_samples DualMesh::samples_
 * (3 per node; general domains only, NULL on boxes). omega receives
 * n_edges values. */
int vsx_advect_circulation(vsx_backend* backend, const double* dual_velocity, const double* node_samples, double dt,
                           double* omega);

/* SimplicialCompute3D::advectDensity: vertex_density is vertexDensity_
 * (n_vertices floats), decay = exp(-smokeDecay * dt) as the solver computes
 * it; out receives n_vertices floats. */
int vsx_advect_density(vsx_backend* backend, const double* dual_velocity, const double* node_samples,
                       const float* vertex_density, double dt, double decay, float* out);

/* SimplicialCompute3D::recover: solve for the edge potential with right-hand
 * side rhs (n_edges). General domains start from phi (the solver's phi_);
 * boxes start from zero and ignore it. On success x holds n_edges values and
 * result the LinearSolveResult fields. */
typedef struct vsx_solve_result {
    int32_t iterations;
    int32_t converged;
    double initial_residual;
    double final_residual;
} vsx_solve_result;

int vsx_recover(vsx_backend* backend, const double* rhs, const double* phi, int32_t max_iterations,
                double absolute_tolerance, double relative_tolerance, double* x, vsx_solve_result* result);

/* Device time of the calling thread's last stage call's kernels, in
 * milliseconds (HIP events around the launches on the backend's stream). */
double vsx_last_kernel_ms(void);

#ifdef __cplusplus
}
#endif

#endif
