// vsx_prepare.hpp - the simplicial backend's mesh, checked and rearranged on
// the host before it goes to the device.
//
// prepare() checks every index of a vsx_mesh against the range of the array
// it indexes (so no kernel reads outside an array), and derives what the
// kernels gather: each face's first vertex, absolute cone plane indices, the
// per-edge and per-vertex incidence in the reference's summation orders and
// the IC(0) level schedules. make_views() then points View / CgOperator at
// the arrays through `map` (an upload to the device, or the host arrays
// themselves in a CPU build).
#pragma once

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/vapor_simplicial3d.h"
#include "vgx_host.hpp"
#include "vsx_device.hpp"

namespace vsx {

struct Prepared {
    int nv = 0, nt = 0, nf = 0, ne = 0, nd = 0, nde = 0, nn = 0, box = 0, bins = 0;
    double size = 0;
    std::vector<double> vert, tetg, face_n, dual_pos, planes, node_pos, sub_grad;
    std::vector<int> tetv, bin_off, bin_items, wall_off, wall_items, face_v0, face_tets;
    std::vector<int> cell_plane_off, cell_corner_off, corner_vertex, corner_cone_off, cone_planes;
    std::vector<int> dual_hint, primal_hint, sub_off, sub_nodes;
    std::vector<int> dual_edges, loop_off, loop_edges, loop_signs;
    // operator
    std::vector<int> face_edges, vertex_off, vertex_items, edge_off, edge_items, edge_vertices;
    std::vector<double> face_signs, star0, star1, star2, diagonal;
    std::vector<unsigned char> vertex_boundary, edge_boundary;
    int forward_levels = 0, backward_levels = 0;
    std::vector<int> forward_off, forward_rows, lower_off, lower_columns, backward_off, backward_rows, upper_off,
        upper_rows;
    std::vector<double> lower_values, upper_values, ic_diagonal;
};

namespace prep {

inline void need(bool ok, const std::string& what) {
    if (!ok) vgx_host::invalid("simplicial mesh: " + what);
}

inline void need_array(const void* p, long count, const char* name) {
    need(count == 0 || p != nullptr, std::string("missing array ") + name);
}

inline void check_range(const int32_t* a, long count, long lo, long hi, const char* name) {
    for (long i = 0; i < count; ++i)
        if (a[i] < lo || a[i] >= hi) need(false, std::string(name) + " index out of range");
}

// CSR offsets of `rows` rows: start at 0, never decrease. Returns the total.
inline long check_offsets(const int32_t* off, long rows, const char* name) {
    need_array(off, rows + 1, name);
    need(off[0] == 0, std::string(name) + " must start at 0");
    for (long r = 0; r < rows; ++r)
        if (off[r + 1] < off[r]) need(false, std::string(name) + " must not decrease");
    return off[rows];
}

inline void check_finite(const double* a, long count, const char* name) {
    for (long i = 0; i < count; ++i)
        if (!std::isfinite(a[i])) need(false, std::string("non-finite ") + name);
}

template <class T> std::vector<T> copy(const T* a, long count) { return count > 0 ? std::vector<T>(a, a + count) : std::vector<T>(); }

// Rows grouped by level, ascending index within a level.
inline void level_schedule(const std::vector<int>& level, std::vector<int>& off, std::vector<int>& rows, int& levels) {
    levels = 0;
    for (int l : level) levels = std::max(levels, l + 1);
    off.assign(levels + 1, 0);
    for (int l : level) ++off[l + 1];
    for (int l = 0; l < levels; ++l) off[l + 1] += off[l];
    rows.assign(level.size(), 0);
    std::vector<int> cursor(off.begin(), off.end() - 1);
    for (int i = 0; i < (int)level.size(); ++i) rows[cursor[level[i]]++] = i;
}

}  // namespace prep

inline Prepared prepare(const vsx_mesh& m) {
    using namespace prep;
    const long nv = m.n_vertices, nt = m.n_tets, nf = m.n_faces, ne = m.n_edges, nd = m.n_dual,
               nde = m.n_dual_edges;
    need(nv > 0 && nt > 0 && nf > 0 && ne > 0 && nd > 0 && nde > 0, "empty mesh");
    need(std::isfinite(m.size) && m.size > 0, "invalid size");
    need(m.bins_per_axis >= 0 && m.bins_per_axis <= 1024, "invalid bins_per_axis");
    need_array(m.vertices, 3 * nv, "vertices");
    check_finite(m.vertices, 3 * nv, "vertices");
    need_array(m.tet_vertices, 4 * nt, "tet_vertices");
    check_range(m.tet_vertices, 4 * nt, 0, nv, "tet_vertices");
    need_array(m.tet_gradients, 9 * nt, "tet_gradients");
    const long nb = (long)m.bins_per_axis * m.bins_per_axis * m.bins_per_axis;
    need_array(m.face_vertices, 3 * nf, "face_vertices");
    check_range(m.face_vertices, 3 * nf, 0, nv, "face_vertices");
    need_array(m.face_edges, 3 * nf, "face_edges");
    check_range(m.face_edges, 3 * nf, 0, ne, "face_edges");
    need_array(m.face_signs, 3 * nf, "face_signs");
    for (long i = 0; i < 3 * nf; ++i) need(m.face_signs[i] == 1 || m.face_signs[i] == -1, "face_signs must be +-1");
    need_array(m.face_tets, 2 * nf, "face_tets");
    check_range(m.face_tets, 2 * nf, -1, nt, "face_tets");
    for (long f = 0; f < nf; ++f) need(m.face_tets[2 * f] >= 0 || m.face_tets[2 * f + 1] >= 0, "face without a tet");
    need_array(m.face_area_normals, 3 * nf, "face_area_normals");
    need_array(m.edge_vertices, 2 * ne, "edge_vertices");
    check_range(m.edge_vertices, 2 * ne, 0, nv, "edge_vertices");
    need_array(m.edge_boundary, ne, "edge_boundary");
    need_array(m.vertex_boundary, nv, "vertex_boundary");
    need_array(m.star0, nv, "star0");
    need_array(m.star1, ne, "star1");
    need_array(m.star2, nf, "star2");
    need_array(m.diagonal, ne, "diagonal");
    need_array(m.dual_positions, 3 * nd, "dual_positions");
    need_array(m.dual_edges, 2 * nde, "dual_edges");
    check_range(m.dual_edges, 2 * nde, 0, nd, "dual_edges");
    const long n_loop_items = check_offsets(m.loop_offsets, ne, "loop_offsets");
    need_array(m.loop_edges, n_loop_items, "loop_edges");
    check_range(m.loop_edges, n_loop_items, 0, nde, "loop_edges");
    need_array(m.loop_signs, n_loop_items, "loop_signs");
    for (long i = 0; i < n_loop_items; ++i) need(m.loop_signs[i] == 1 || m.loop_signs[i] == -1, "loop_signs must be +-1");

    Prepared P;
    P.nv = (int)nv;
    P.nt = (int)nt;
    P.nf = (int)nf;
    P.ne = (int)ne;
    P.nd = (int)nd;
    P.nde = (int)nde;
    P.box = m.box_domain != 0 ? 1 : 0;
    P.size = m.size;
    P.bins = (int)m.bins_per_axis;
    P.vert = copy(m.vertices, 3 * nv);
    P.tetv = copy(m.tet_vertices, 4 * nt);
    P.tetg = copy(m.tet_gradients, 9 * nt);
    if (nb > 0) {
        const long n_bin_items = check_offsets(m.bin_offsets, nb, "bin_offsets");
        need_array(m.bin_items, n_bin_items, "bin_items");
        check_range(m.bin_items, n_bin_items, 0, nt, "bin_items");
        const long n_wall_items = check_offsets(m.wall_offsets, nb, "wall_offsets");
        need_array(m.wall_items, n_wall_items, "wall_items");
        check_range(m.wall_items, n_wall_items, 0, nf, "wall_items");
        P.bin_off = copy(m.bin_offsets, nb + 1);
        P.bin_items = copy(m.bin_items, n_bin_items);
        P.wall_off = copy(m.wall_offsets, nb + 1);
        P.wall_items = copy(m.wall_items, n_wall_items);
    }
    P.face_v0.resize(nf);
    for (long f = 0; f < nf; ++f) P.face_v0[f] = m.face_vertices[3 * f];
    P.face_tets = copy(m.face_tets, 2 * nf);
    P.face_n = copy(m.face_area_normals, 3 * nf);
    P.dual_pos = copy(m.dual_positions, 3 * nd);
    P.dual_edges = copy(m.dual_edges, 2 * nde);
    P.loop_off = copy(m.loop_offsets, ne + 1);
    P.loop_edges = copy(m.loop_edges, n_loop_items);
    P.loop_signs = copy(m.loop_signs, n_loop_items);

    if (P.box) {
        const long n_planes = check_offsets(m.cell_plane_offsets, nv, "cell_plane_offsets");
        need_array(m.cell_planes, 4 * n_planes, "cell_planes");
        const long n_corners = check_offsets(m.cell_corner_offsets, nv, "cell_corner_offsets");
        need_array(m.corner_vertices, n_corners, "corner_vertices");
        check_range(m.corner_vertices, n_corners, 0, nd, "corner_vertices");
        const long n_cones = check_offsets(m.corner_cone_offsets, n_corners, "corner_cone_offsets");
        need_array(m.cone_planes, 3 * n_cones, "cone_planes");
        // Plane indices are local to the cell; the kernels take them absolute.
        P.cone_planes.assign(3 * n_cones, 0);
        for (long c = 0; c < nv; ++c) {
            const int base = m.cell_plane_offsets[c], count = m.cell_plane_offsets[c + 1] - base;
            for (long k = m.cell_corner_offsets[c]; k < m.cell_corner_offsets[c + 1]; ++k)
                for (long j = 3L * m.corner_cone_offsets[k]; j < 3L * m.corner_cone_offsets[k + 1]; ++j) {
                    need(m.cone_planes[j] >= 0 && m.cone_planes[j] < count, "cone_planes index out of range");
                    P.cone_planes[j] = base + m.cone_planes[j];
                }
        }
        P.cell_plane_off = copy(m.cell_plane_offsets, nv + 1);
        P.planes = copy(m.cell_planes, 4 * n_planes);
        P.cell_corner_off = copy(m.cell_corner_offsets, nv + 1);
        P.corner_vertex = copy(m.corner_vertices, n_corners);
        P.corner_cone_off = copy(m.corner_cone_offsets, n_corners + 1);
    } else {
        const long nn = m.n_nodes;
        need(nn >= nv, "general domains need a subdivision node per primal vertex");
        need_array(m.dual_hints, nd, "dual_hints");
        need_array(m.primal_hints, nv, "primal_hints");
        need_array(m.node_positions, 3 * nn, "node_positions");
        const long n_subs = check_offsets(m.sub_offsets, nt, "sub_offsets");
        need_array(m.sub_nodes, 4 * n_subs, "sub_nodes");
        check_range(m.sub_nodes, 4 * n_subs, 0, nn, "sub_nodes");
        need_array(m.sub_gradients, 9 * n_subs, "sub_gradients");
        P.nn = (int)nn;
        P.dual_hint = copy(m.dual_hints, nd);
        P.primal_hint = copy(m.primal_hints, nv);
        P.node_pos = copy(m.node_positions, 3 * nn);
        P.sub_off = copy(m.sub_offsets, nt + 1);
        P.sub_nodes = copy(m.sub_nodes, 4 * n_subs);
        P.sub_grad = copy(m.sub_gradients, 9 * n_subs);
    }

    // The operator, gathered in the reference's summation orders.
    P.face_edges = copy(m.face_edges, 3 * nf);
    P.face_signs.resize(3 * nf);
    for (long i = 0; i < 3 * nf; ++i) P.face_signs[i] = (double)m.face_signs[i];
    P.star2 = copy(m.star2, nf);
    P.star1 = copy(m.star1, ne);
    P.star0 = copy(m.star0, nv);
    P.vertex_boundary = copy(m.vertex_boundary, nv);
    P.edge_boundary = copy(m.edge_boundary, ne);
    P.edge_vertices = copy(m.edge_vertices, 2 * ne);
    P.diagonal = copy(m.diagonal, ne);
    {
        // Faces per edge in ascending face order (Mesh::d1Transpose).
        P.edge_off.assign(ne + 1, 0);
        P.edge_items.assign(3 * nf, 0);
        for (long i = 0; i < 3 * nf; ++i) ++P.edge_off[m.face_edges[i] + 1];
        for (long e = 0; e < ne; ++e) P.edge_off[e + 1] += P.edge_off[e];
        std::vector<int> cursor(P.edge_off.begin(), P.edge_off.end() - 1);
        for (long f = 0; f < nf; ++f)
            for (int k = 0; k < 3; ++k) {
                const int e = m.face_edges[3 * f + k];
                P.edge_items[cursor[e]++] = m.face_signs[3 * f + k] > 0 ? (int)f : ~(int)f;
            }
    }
    {
        // Edges per vertex in ascending edge order (the divergence loop).
        P.vertex_off.assign(nv + 1, 0);
        P.vertex_items.assign(2 * ne, 0);
        for (long i = 0; i < 2 * ne; ++i) ++P.vertex_off[m.edge_vertices[i] + 1];
        for (long v = 0; v < nv; ++v) P.vertex_off[v + 1] += P.vertex_off[v];
        std::vector<int> cursor(P.vertex_off.begin(), P.vertex_off.end() - 1);
        for (long e = 0; e < ne; ++e) {
            P.vertex_items[cursor[m.edge_vertices[2 * e]]++] = ~(int)e;
            P.vertex_items[cursor[m.edge_vertices[2 * e + 1]]++] = (int)e;
        }
    }
    if (m.ic_offsets) {
        const long nnz = check_offsets(m.ic_offsets, ne, "ic_offsets");
        need_array(m.ic_columns, nnz, "ic_columns");
        need_array(m.ic_values, nnz, "ic_values");
        need_array(m.ic_diagonal, ne, "ic_diagonal");
        for (long i = 0; i < ne; ++i)
            for (long j = m.ic_offsets[i]; j < m.ic_offsets[i + 1]; ++j) {
                need(m.ic_columns[j] >= 0 && m.ic_columns[j] < i, "ic_columns must lie below the diagonal");
                need(j == m.ic_offsets[i] || m.ic_columns[j] > m.ic_columns[j - 1], "ic_columns must ascend");
            }
        std::vector<int> level(ne, 0);
        for (long i = 0; i < ne; ++i)
            for (long j = m.ic_offsets[i]; j < m.ic_offsets[i + 1]; ++j)
                level[i] = std::max(level[i], level[m.ic_columns[j]] + 1);
        level_schedule(level, P.forward_off, P.forward_rows, P.forward_levels);
        // Upper rows: for row j, the rows i > j whose lower row holds j, in
        // descending i (the backward sweep's order of updates to z[j]).
        P.upper_off.assign(ne + 1, 0);
        P.upper_rows.assign(nnz, 0);
        P.upper_values.assign(nnz, 0);
        for (long j = 0; j < nnz; ++j) ++P.upper_off[m.ic_columns[j] + 1];
        for (long r = 0; r < ne; ++r) P.upper_off[r + 1] += P.upper_off[r];
        std::vector<int> cursor(P.upper_off.begin(), P.upper_off.end() - 1);
        for (long i = ne - 1; i >= 0; --i)
            for (long j = m.ic_offsets[i]; j < m.ic_offsets[i + 1]; ++j) {
                const int c = m.ic_columns[j];
                P.upper_rows[cursor[c]] = (int)i;
                P.upper_values[cursor[c]++] = m.ic_values[j];
            }
        std::fill(level.begin(), level.end(), 0);
        for (long i = ne - 1; i >= 0; --i)
            for (long j = m.ic_offsets[i]; j < m.ic_offsets[i + 1]; ++j)
                level[m.ic_columns[j]] = std::max(level[m.ic_columns[j]], level[i] + 1);
        level_schedule(level, P.backward_off, P.backward_rows, P.backward_levels);
        P.lower_off = copy(m.ic_offsets, ne + 1);
        P.lower_columns = copy(m.ic_columns, nnz);
        P.lower_values = copy(m.ic_values, nnz);
        P.ic_diagonal = copy(m.ic_diagonal, ne);
    }
    return P;
}

// View and CgOperator over the arrays `map` returns for each vector.
template <class Map> void make_views(const Prepared& P, Map&& map, View& V, CgOperator& A) {
    V = View{};
    V.nv = P.nv;
    V.nt = P.nt;
    V.nf = P.nf;
    V.ne = P.ne;
    V.nd = P.nd;
    V.nn = P.nn;
    V.box = P.box;
    V.size = P.size;
    V.bins = P.bins;
    V.vert = map(P.vert);
    V.tetv = map(P.tetv);
    V.tetg = map(P.tetg);
    V.bin_off = map(P.bin_off);
    V.bin_items = map(P.bin_items);
    V.wall_off = map(P.wall_off);
    V.wall_items = map(P.wall_items);
    V.face_v0 = map(P.face_v0);
    V.face_tets = map(P.face_tets);
    V.face_n = map(P.face_n);
    V.cell_plane_off = map(P.cell_plane_off);
    V.planes = map(P.planes);
    V.cell_corner_off = map(P.cell_corner_off);
    V.corner_vertex = map(P.corner_vertex);
    V.corner_cone_off = map(P.corner_cone_off);
    V.cone_planes = map(P.cone_planes);
    V.dual_pos = map(P.dual_pos);
    V.dual_hint = map(P.dual_hint);
    V.primal_hint = map(P.primal_hint);
    V.node_pos = map(P.node_pos);
    V.sub_off = map(P.sub_off);
    V.sub_nodes = map(P.sub_nodes);
    V.sub_grad = map(P.sub_grad);

    A = CgOperator{};
    A.ne = P.ne;
    A.nf = P.nf;
    A.nv = P.nv;
    A.face_edges = map(P.face_edges);
    A.face_signs = map(P.face_signs);
    A.star2 = map(P.star2);
    A.vertex_off = map(P.vertex_off);
    A.vertex_items = map(P.vertex_items);
    A.star1 = map(P.star1);
    A.star0 = map(P.star0);
    A.vertex_boundary = map(P.vertex_boundary);
    A.edge_off = map(P.edge_off);
    A.edge_items = map(P.edge_items);
    A.edge_vertices = map(P.edge_vertices);
    A.edge_boundary = map(P.edge_boundary);
    A.diagonal = map(P.diagonal);
    A.forward_levels = P.forward_levels;
    A.backward_levels = P.backward_levels;
    A.forward_off = map(P.forward_off);
    A.forward_rows = map(P.forward_rows);
    A.lower_off = map(P.lower_off);
    A.lower_columns = map(P.lower_columns);
    A.lower_values = map(P.lower_values);
    A.backward_off = map(P.backward_off);
    A.backward_rows = map(P.backward_rows);
    A.upper_off = map(P.upper_off);
    A.upper_rows = map(P.upper_rows);
    A.upper_values = map(P.upper_values);
    A.ic_diagonal = map(P.ic_diagonal);
}

}  // namespace vsx
