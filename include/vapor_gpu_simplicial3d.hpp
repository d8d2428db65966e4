// vapor_gpu_simplicial3d.hpp - vapor's GpuSimplicial3D backend on the HIP
// library (vapor_simplicial3d.h), for vapor's own tree.
//
// vapor declares GpuSimplicial3D in
// src/solvers/simplicial3d/gpu/GpuSimplicial3D.hpp:5-19 and implements it over
// OpenGL compute shaders in gpu/GpuSimplicial3D.cpp. This header implements
// the same declaration over the HIP library instead: a maintainer compiles it
// in place of gpu/GpuSimplicial3D.cpp (include it from one source file, or
// from several: every definition is inline) and links libvapor_mac3d.so.
// The solver, its mesh and its dual name GpuSimplicial3D a friend
// (SimplicialFluidSolver3D.hpp:80, SimplicialMesh3D.hpp:52,
// SimplicialDual3D.hpp:36); initialize() reads their private arrays through
// that friendship, as the OpenGL backend does, and hands them to
// vsx_backend_create once per mesh.
//
// Every stage call keeps the reference's contract: advectCirculation fills
// omega with one value per primal edge, advectDensity one float per primal
// vertex, recover returns LinearSolveResult and the potential. Errors become
// the reference's exceptions: std::invalid_argument for bad arguments and
// std::runtime_error with the CPU solver's message for failures inside a
// stage.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "solvers/simplicial3d/SimplicialFluidSolver3D.hpp"
#include "solvers/simplicial3d/gpu/GpuSimplicial3D.hpp"
#include "vapor_simplicial3d.h"

namespace vapor_hip_simplicial {

inline void raise(int status) {
    if (status == VGX_OK) return;
    const std::string message = vgx_last_error();
    if (status == VGX_EINVAL) throw std::invalid_argument(message);
    if (status == VGX_ERANGE) throw std::length_error(message);
    throw std::runtime_error(message);
}

static_assert(sizeof(simplicial3d::Point) == 3 * sizeof(double), "Point must be three packed doubles");

inline const double* points(const std::vector<simplicial3d::Point>& v) {
    return v.empty() ? nullptr : &v.front().x;
}

// The state a stage call reads: one velocity per dual vertex and, on general
// domains, one subdivision sample per node.
inline void check_state(const SimplicialFluidSolver3D& solver, const std::vector<simplicial3d::Point>& velocity,
                        const std::vector<simplicial3d::Point>& samples, size_t nodes) {
    if (velocity.size() != solver.dual().vertices.size()) throw std::invalid_argument("Dual velocity size mismatch");
    if (!solver.domain().boxDomain && samples.size() != nodes)
        throw std::invalid_argument("Subdivision sample size mismatch");
}

}  // namespace vapor_hip_simplicial

struct GpuSimplicial3D::Impl {
    vsx_backend* backend = nullptr;
    const simplicial3d::Tet* meshKey = nullptr;
    double size = 0;
    size_t edges = 0, duals = 0;
    std::string device;
    ~Impl() { vsx_backend_destroy(backend); }
};

inline GpuSimplicial3D::GpuSimplicial3D() : impl_(std::make_unique<Impl>()) {}
inline GpuSimplicial3D::~GpuSimplicial3D() = default;
inline const std::string& GpuSimplicial3D::device() const { return impl_->device; }

// GpuSimplicial3D::initialize (gpu/GpuSimplicial3D.cpp:54-108): flatten the
// mesh, the dual, the bins, the box cells or the subdivision, and the
// operator (the Jacobi diagonal and, on general domains, the IC(0) factor),
// once per mesh.
inline void GpuSimplicial3D::initialize(const SimplicialFluidSolver3D& solver) {
    using namespace simplicial3d;
    auto& gpu = *impl_;
    const Mesh& mesh = solver.mesh_;
    const DualMesh& dual = solver.dual_;
    if (gpu.backend && gpu.meshKey == mesh.tets.data() && gpu.size == mesh.size &&
        gpu.edges == mesh.edges.size() && gpu.duals == dual.vertices.size())
        return;
    auto point = [](std::vector<double>& data, Point p) { data.insert(data.end(), {p.x, p.y, p.z}); };
    auto csr = [](const auto& rows, std::vector<int32_t>& offsets, std::vector<int32_t>& items) {
        offsets.assign(1, 0);
        items.clear();
        for (const auto& row : rows) {
            items.insert(items.end(), row.begin(), row.end());
            offsets.push_back(int32_t(items.size()));
        }
    };
    vsx_mesh m{};
    m.n_vertices = int32_t(mesh.vertices.size());
    m.n_tets = int32_t(mesh.tets.size());
    m.n_faces = int32_t(mesh.faces.size());
    m.n_edges = int32_t(mesh.edges.size());
    m.box_domain = mesh.boxDomain ? 1 : 0;
    m.size = mesh.size;

    std::vector<double> vertices, tetGradients;
    std::vector<int32_t> tetVertices;
    for (auto p : mesh.vertices) point(vertices, p);
    for (const auto& t : mesh.tets) {
        tetVertices.insert(tetVertices.end(), t.vertices.begin(), t.vertices.end());
        for (int k = 1; k < 4; ++k) point(tetGradients, t.gradients[k]);
    }
    m.vertices = vertices.data();
    m.tet_vertices = tetVertices.data();
    m.tet_gradients = tetGradients.data();

    std::vector<int32_t> binOffsets, binItems, wallOffsets, wallItems;
    m.bins_per_axis = mesh.binsPerAxis_;
    if (mesh.binsPerAxis_ > 0) {
        csr(mesh.bins_, binOffsets, binItems);
        csr(mesh.boundaryBins_, wallOffsets, wallItems);
        m.bin_offsets = binOffsets.data();
        m.bin_items = binItems.data();
        m.wall_offsets = wallOffsets.data();
        m.wall_items = wallItems.data();
    }

    std::vector<int32_t> faceVertices, faceEdges, faceSigns, faceTets, edgeVertices;
    std::vector<double> areaNormals;
    for (const auto& f : mesh.faces) {
        faceVertices.insert(faceVertices.end(), f.vertices.begin(), f.vertices.end());
        faceEdges.insert(faceEdges.end(), f.edges.begin(), f.edges.end());
        faceSigns.insert(faceSigns.end(), f.signs.begin(), f.signs.end());
        faceTets.insert(faceTets.end(), f.tets.begin(), f.tets.end());
        point(areaNormals, f.areaNormal);
    }
    std::vector<uint8_t> edgeBoundary, vertexBoundary;
    for (const auto& e : mesh.edges) {
        edgeVertices.insert(edgeVertices.end(), e.vertices.begin(), e.vertices.end());
        edgeBoundary.push_back(e.boundary ? 1 : 0);
    }
    for (bool b : mesh.boundaryVertices) vertexBoundary.push_back(b ? 1 : 0);
    m.face_vertices = faceVertices.data();
    m.face_edges = faceEdges.data();
    m.face_signs = faceSigns.data();
    m.face_tets = faceTets.data();
    m.face_area_normals = areaNormals.data();
    m.edge_vertices = edgeVertices.data();
    m.edge_boundary = edgeBoundary.data();
    m.vertex_boundary = vertexBoundary.data();
    m.star0 = mesh.star0.data();
    m.star1 = mesh.star1.data();
    m.star2 = mesh.star2.data();
    m.diagonal = solver.diagonal_.data();

    std::vector<int32_t> icOffsets, icColumns;
    std::vector<double> icValues;
    if (!solver.incompleteLower_.empty()) {
        icOffsets.push_back(0);
        for (const auto& row : solver.incompleteLower_) {
            for (auto [j, value] : row) {
                icColumns.push_back(j);
                icValues.push_back(value);
            }
            icOffsets.push_back(int32_t(icColumns.size()));
        }
        m.ic_offsets = icOffsets.data();
        m.ic_columns = icColumns.data();
        m.ic_values = icValues.data();
        m.ic_diagonal = solver.incompleteDiagonal_.data();
    }

    std::vector<double> dualPositions;
    std::vector<int32_t> dualEdges, loopOffsets{0}, loopEdges, loopSigns;
    for (const auto& v : dual.vertices) point(dualPositions, v.position);
    for (auto e : dual.edges) dualEdges.insert(dualEdges.end(), {e.a, e.b});
    for (const auto& loop : dual.loops) {
        for (auto side : loop) {
            loopEdges.push_back(side.edge);
            loopSigns.push_back(side.sign);
        }
        loopOffsets.push_back(int32_t(loopEdges.size()));
    }
    m.n_dual = int32_t(dual.vertices.size());
    m.n_dual_edges = int32_t(dual.edges.size());
    m.dual_positions = dualPositions.data();
    m.dual_edges = dualEdges.data();
    m.loop_offsets = loopOffsets.data();
    m.loop_edges = loopEdges.data();
    m.loop_signs = loopSigns.data();

    std::vector<int32_t> planeOffsets{0}, cornerOffsets{0}, cornerVertices, coneOffsets{0}, conePlanes;
    std::vector<double> planes;
    std::vector<int32_t> subOffsets{0}, subNodes;
    std::vector<double> nodePositions, subGradients;
    if (mesh.boxDomain) {
        for (const auto& cell : dual.cells) {
            for (const auto& plane : cell.planes) {
                point(planes, plane.normal);
                planes.push_back(plane.offset);
            }
            planeOffsets.push_back(int32_t(planes.size() / 4));
            for (const auto& corner : cell.corners) {
                cornerVertices.push_back(corner.vertex);
                for (const auto& cone : corner.cones) conePlanes.insert(conePlanes.end(), cone.begin(), cone.end());
                coneOffsets.push_back(int32_t(conePlanes.size() / 3));
            }
            cornerOffsets.push_back(int32_t(cornerVertices.size()));
        }
        if (dual.cells.size() != mesh.vertices.size())
            throw std::invalid_argument("Box dual needs one cell per primal vertex");
        m.cell_plane_offsets = planeOffsets.data();
        m.cell_planes = planes.data();
        m.cell_corner_offsets = cornerOffsets.data();
        m.corner_vertices = cornerVertices.data();
        m.corner_cone_offsets = coneOffsets.data();
        m.cone_planes = conePlanes.data();
    } else {
        for (const auto& node : dual.nodes_) point(nodePositions, node.position);
        for (const auto& subdivision : dual.subdivisions_) {
            for (const auto& sub : subdivision) {
                subNodes.insert(subNodes.end(), sub.nodes.begin(), sub.nodes.end());
                for (auto g : sub.gradients) point(subGradients, g);
            }
            subOffsets.push_back(int32_t(subNodes.size() / 4));
        }
        if (dual.subdivisions_.size() != mesh.tets.size())
            throw std::invalid_argument("General domains need one subdivision per tet");
        m.n_nodes = int32_t(dual.nodes_.size());
        m.dual_hints = solver.dualTets_.data();
        m.primal_hints = solver.primalTets_.data();
        m.node_positions = nodePositions.data();
        m.sub_offsets = subOffsets.data();
        m.sub_nodes = subNodes.data();
        m.sub_gradients = subGradients.data();
    }

    vsx_backend* created = nullptr;
    vapor_hip_simplicial::raise(vsx_backend_create(&m, &created));
    vsx_backend_destroy(gpu.backend);
    gpu.backend = created;
    gpu.device = vsx_backend_device(created);
    gpu.meshKey = mesh.tets.data();
    gpu.size = mesh.size;
    gpu.edges = mesh.edges.size();
    gpu.duals = dual.vertices.size();
}

inline void GpuSimplicial3D::advectCirculation(const SimplicialFluidSolver3D& solver, double dt,
                                               std::vector<double>& omega) {
    initialize(solver);
    vapor_hip_simplicial::check_state(solver, solver.dualVelocity_, solver.dual_.samples_, solver.dual_.nodes_.size());
    std::vector<double> next(solver.mesh_.edges.size());
    vapor_hip_simplicial::raise(vsx_advect_circulation(impl_->backend,
                                                       vapor_hip_simplicial::points(solver.dualVelocity_),
                                                       vapor_hip_simplicial::points(solver.dual_.samples_), dt,
                                                       next.data()));
    omega = std::move(next);
}

inline void GpuSimplicial3D::advectDensity(const SimplicialFluidSolver3D& solver, double dt,
                                           std::vector<float>& density) {
    initialize(solver);
    vapor_hip_simplicial::check_state(solver, solver.dualVelocity_, solver.dual_.samples_, solver.dual_.nodes_.size());
    if (solver.vertexDensity_.size() != solver.mesh_.vertices.size())
        throw std::invalid_argument("Density size mismatch");
    // The decay as SimplicialFluidSolver3D.cpp:597 computes it.
    const double decay = std::exp(-solver.parameters_.smokeDecay * dt);
    std::vector<float> next(solver.mesh_.vertices.size());
    vapor_hip_simplicial::raise(vsx_advect_density(impl_->backend, vapor_hip_simplicial::points(solver.dualVelocity_),
                                                   vapor_hip_simplicial::points(solver.dual_.samples_),
                                                   solver.vertexDensity_.data(), dt, decay, next.data()));
    density = std::move(next);
}

inline LinearSolveResult GpuSimplicial3D::recover(const SimplicialFluidSolver3D& solver, std::span<const double> rhs,
                                                  std::vector<double>& x) {
    initialize(solver);
    const size_t n = solver.mesh_.edges.size();
    if (rhs.size() != n || solver.phi_.size() != n) throw std::invalid_argument("Potential size mismatch");
    const auto& p = solver.parameters_;
    std::vector<double> next(n);
    vsx_solve_result r{};
    vapor_hip_simplicial::raise(vsx_recover(impl_->backend, rhs.data(), solver.phi_.data(), p.cgIterations,
                                            p.cgAbsoluteTolerance, p.cgRelativeTolerance, next.data(), &r));
    x = std::move(next);
    LinearSolveResult result;
    result.iterations = r.iterations;
    result.initialResidual = r.initial_residual;
    result.finalResidual = r.final_residual;
    result.converged = r.converged != 0;
    return result;
}
