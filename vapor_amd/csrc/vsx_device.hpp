// vsx_device.hpp - geometry of vapor's simplicial 3D solver, for the device
// and the host.
//
// Each function restates one routine of src/solvers/simplicial3d with the
// same operations in the same order (the library is compiled with
// -ffp-contract=off), so a traced point, a weight or a sample rounds as the
// reference's CPU code rounds it. The reference's comparisons are kept as
// written, including the ones that decide which of two equal values wins
// (std::min_element, std::max, a strict '<' in a nearest search).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vsx {

// Error bits a kernel raises; the host turns them into the reference's
// exceptions.
enum : int {
    kErrWeights = 1,      // "Invalid generalized barycentric weights"
    kErrSubdivision = 2,  // "Point outside circumcentric subdivision"
    kErrNonFiniteRhs = 4, // "Non-finite CG RHS"
    kErrDefinite = 8,     // "CG lost positive definiteness"
};

struct P3 {
    double x, y, z;
};
// simplicial3d::Point arithmetic and dot / cross / length
// (SimplicialMesh3D.hpp:8-16, SimplicialMesh3D.cpp:11).
__host__ __device__ __forceinline__ P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ P3 mul(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ P3 cross(P3 a, P3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ __forceinline__ double length(P3 a) { return __builtin_sqrt(dot(a, a)); }
__host__ __device__ __forceinline__ P3 ld3(const double* a, long i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
__host__ __device__ __forceinline__ void st3(double* a, long i, P3 v) {
    a[3 * i] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
}

// std::max(a, b), std::min(a, b), std::clamp(v, lo, hi) as libstdc++ defines them.
__host__ __device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double sclamp(double v, double lo, double hi) {
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}
// *std::min_element over four weights: a later value replaces the current one
// only if it compares smaller.
__host__ __device__ __forceinline__ double min4(const double w[4]) {
    double m = w[0];
    if (w[1] < m) m = w[1];
    if (w[2] < m) m = w[2];
    if (w[3] < m) m = w[3];
    return m;
}
// int(v) for the bin of a coordinate, made total: NaN and out-of-range values
// go to bin 0 or the last bin. Equal to the reference's int(v) clamped to
// [0, n-1] wherever int(v) is defined.
__host__ __device__ __forceinline__ int bin_index(double v, int n) {
    if (!(v >= 0.0)) return 0;
    if (v >= (double)n) return n - 1;
    const int b = (int)v;
    return b < n - 1 ? b : n - 1;
}

// The mesh as the kernels read it (device pointers; host pointers in a CPU
// build).
struct View {
    int nv, nt, nf, ne, nd, nn;
    int box;
    double size;
    const double* vert;   // 3 per vertex
    const int* tetv;      // 4 per tet
    const double* tetg;   // 9 per tet: gradients 1..3
    int bins;
    const int* bin_off;
    const int* bin_items;
    const int* wall_off;
    const int* wall_items;
    const int* face_v0;   // Face::vertices[0]
    const int* face_tets; // 2 per face
    const double* face_n; // 3 per face
    // box cells
    const int* cell_plane_off;
    const double* planes; // 4 per plane
    const int* cell_corner_off;
    const int* corner_vertex;
    const int* corner_cone_off;
    const int* cone_planes; // 3 per cone, absolute plane index
    // dual and subdivision
    const double* dual_pos;
    const int* dual_hint;
    const int* primal_hint;
    const double* node_pos;
    const int* sub_off;
    const int* sub_nodes;
    const double* sub_grad; // 9 per subtet
};

// Mesh::barycentric (SimplicialMesh3D.cpp:166-170).
__host__ __device__ __forceinline__ void barycentric(const View& V, int t, P3 p, double w[4]) {
    const P3 q = sub(p, ld3(V.vert, V.tetv[4 * t]));
    w[1] = dot(ld3(V.tetg, 3L * t), q);
    w[2] = dot(ld3(V.tetg, 3L * t + 1), q);
    w[3] = dot(ld3(V.tetg, 3L * t + 2), q);
    w[0] = 1 - w[1] - w[2] - w[3];
}

__host__ __device__ __forceinline__ bool inside(const View& V, int t, P3 p, double tolerance) {
    double w[4];
    barycentric(V, t, p, w);
    return min4(w) >= tolerance;
}

// Mesh::locate (SimplicialMesh3D.cpp:171-183): the hint, then the point's
// bin in list order, or every tet from the hint on when there are no bins.
__host__ __device__ inline int locate(const View& V, P3 p, int hint) {
    if (hint >= 0 && hint < V.nt && inside(V, hint, p, -1e-9)) return hint;
    if (V.bins > 0) {
        if (p.x < 0 || p.y < 0 || p.z < 0 || p.x > V.size || p.y > V.size || p.z > V.size) return -1;
        const int n = V.bins;
        const int x = bin_index(p.x / V.size * n, n), y = bin_index(p.y / V.size * n, n),
                  z = bin_index(p.z / V.size * n, n);
        const int b = x + n * (y + n * z);
        for (int j = V.bin_off[b]; j < V.bin_off[b + 1]; ++j) {
            const int t = V.bin_items[j];
            if (inside(V, t, p, -1e-9)) return t;
        }
        return -1;
    }
    const int start = hint > 0 ? hint : 0;
    for (int k = 0; k < V.nt; ++k) {
        const int t = (int)(((long)start + k) % V.nt);
        if (inside(V, t, p, -1e-9)) return t;
    }
    return -1;
}

// The box trace's clip (SimplicialFluidSolver3D.cpp:418).
__host__ __device__ __forceinline__ P3 clamp_box(P3 q, double size) {
    return {sclamp(q.x, 0., size), sclamp(q.y, 0., size), sclamp(q.z, 0., size)};
}

// Mesh::clipSegment (SimplicialMesh3D.cpp:184-213): the first boundary face
// the segment crosses, found through the wall bins in the reference's order.
__host__ __device__ inline P3 clip_segment(const View& V, P3 start, P3 end, int hint) {
    if (hint >= 0 && hint < V.nt && inside(V, hint, start, -1e-10) && inside(V, hint, end, -1e-10)) return end;
    const P3 delta = sub(end, start);
    if (length(delta) < 1e-14 * V.size) return start;
    double first = 1;
    auto inspect = [&](int fi) {
        const int t0 = V.face_tets[2 * fi], t1 = V.face_tets[2 * fi + 1];
        const int ti = t0 >= 0 ? t0 : t1;
        const P3 normal = mul(ld3(V.face_n, fi), t0 >= 0 ? 1. : -1.);
        const double denominator = dot(normal, delta);
        if (denominator <= 1e-15 * length(normal) * length(delta)) return;
        const double hit = dot(normal, sub(ld3(V.vert, V.face_v0[fi]), start)) / denominator;
        if (hit < -1e-9 || hit > first) return;
        double w[4];
        barycentric(V, ti, add(start, mul(delta, hit)), w);
        if (min4(w) < -1e-8) return;
        first = smax(0., hit);
    };
    if (V.bins > 0) {
        const int n = V.bins;
        auto bin = [&](double c) { return bin_index(c / V.size * n, n); };
        const int z0 = bin(smin(start.z, end.z)), z1 = bin(smax(start.z, end.z));
        const int y0 = bin(smin(start.y, end.y)), y1 = bin(smax(start.y, end.y));
        const int x0 = bin(smin(start.x, end.x)), x1 = bin(smax(start.x, end.x));
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    const int b = x + n * (y + n * z);
                    for (int j = V.wall_off[b]; j < V.wall_off[b + 1]; ++j) inspect(V.wall_items[j]);
                }
    } else {
        for (int fi = 0; fi < V.nf; ++fi)
            if (V.face_tets[2 * fi] < 0 || V.face_tets[2 * fi + 1] < 0) inspect(fi);
    }
    return first < 1 ? add(start, mul(delta, smax(0., first - 1e-8))) : end;
}

// One step of DualMesh::interpolate's nearest-site search
// (SimplicialDual3D.cpp:155-156): a site replaces the current one only if it
// is strictly nearer, so the first nearest site in index order wins.
__host__ __device__ __forceinline__ void nearer(P3 p, P3 site_position, int site, double& best, int& best_site) {
    const P3 d = sub(p, site_position);
    const double squared = dot(d, d);
    if (squared < best) {
        best_site = site;
        best = squared;
    }
}

// The Warren et al. weight of one corner of a box cell
// (SimplicialDual3D.cpp:129-135): one term per cone of the corner's normal
// cone, each plane's distance floored at 1e-13 size.
__host__ __device__ __forceinline__ double corner_weight(const View& V, int corner, P3 p) {
    double weight = 0;
    const double floor = 1e-13 * V.size;
    for (int k = V.corner_cone_off[corner]; k < V.corner_cone_off[corner + 1]; ++k) {
        const int ia = V.cone_planes[3 * k], ib = V.cone_planes[3 * k + 1], ic = V.cone_planes[3 * k + 2];
        const P3 a = ld3(V.planes + 4L * ia, 0), b = ld3(V.planes + 4L * ib, 0), c = ld3(V.planes + 4L * ic, 0);
        const double da = smax(floor, V.planes[4L * ia + 3] - dot(a, p));
        const double db = smax(floor, V.planes[4L * ib + 3] - dot(b, p));
        const double dc = smax(floor, V.planes[4L * ic + 3] - dot(c, p));
        weight += __builtin_fabs(dot(a, cross(b, c))) / (da * db * dc);
    }
    return weight;
}

// DualMesh::coordinates then the weighted sum of DualMesh::interpolate
// (SimplicialDual3D.cpp:124-140, 157-159) over the cell of `site`: a corner
// within 1e-12 size of p takes weight 1 and every other corner 0; otherwise
// the weights are normalised by their sum in corner order. Sets kErrWeights
// where the sum is not finite and positive.
__host__ __device__ inline P3 warren(const View& V, int site, P3 p, const double* values, int& error) {
    const int c0 = V.cell_corner_off[site], c1 = V.cell_corner_off[site + 1];
    int snap = -1;
    for (int c = c0; c < c1; ++c)
        if (length(sub(ld3(V.dual_pos, V.corner_vertex[c]), p)) < 1e-12 * V.size) {
            snap = c;
            break;
        }
    P3 value{0, 0, 0};
    if (snap >= 0) {
        for (int c = c0; c < c1; ++c) value = add(value, mul(ld3(values, V.corner_vertex[c]), c == snap ? 1. : 0.));
        return value;
    }
    double sum = 0;
    for (int c = c0; c < c1; ++c) sum += corner_weight(V, c, p);
    if (!(sum - sum == 0) || sum <= 0) {
        error |= kErrWeights;
        return value;
    }
    for (int c = c0; c < c1; ++c) value = add(value, mul(ld3(values, V.corner_vertex[c]), corner_weight(V, c, p) / sum));
    return value;
}

// DualMesh::interpolate on a general domain (SimplicialDual3D.cpp:143-153):
// the first subtet of the located tet that holds p within -1e-8; zero
// outside the mesh; kErrSubdivision where no subtet holds it.
__host__ __device__ inline P3 subdivision_value(const View& V, P3 p, int hint, const double* samples, int& error) {
    const int t = locate(V, p, hint);
    if (t < 0) return {0, 0, 0};
    for (int s = V.sub_off[t]; s < V.sub_off[t + 1]; ++s) {
        const P3 delta = sub(p, ld3(V.node_pos, V.sub_nodes[4 * s]));
        double w[4];
        w[1] = dot(ld3(V.sub_grad, 3L * s), delta);
        w[2] = dot(ld3(V.sub_grad, 3L * s + 1), delta);
        w[3] = dot(ld3(V.sub_grad, 3L * s + 2), delta);
        w[0] = 1 - w[1] - w[2] - w[3];
        if (min4(w) >= -1e-8) {
            P3 result{0, 0, 0};
            for (int k = 0; k < 4; ++k) result = add(result, mul(ld3(samples, V.sub_nodes[4 * s + k]), w[k]));
            return result;
        }
    }
    error |= kErrSubdivision;
    return {0, 0, 0};
}

// The P1 smoke at q in tet t with clamped, renormalised weights
// (SimplicialFluidSolver3D.cpp:500-511 and 605-608). Returns value / total
// and the total through `total`.
__host__ __device__ inline double tet_density(const View& V, int t, P3 q, const float* density, double& total) {
    double w[4];
    barycentric(V, t, q, w);
    double value = 0;
    total = 0;
    for (int k = 0; k < 4; ++k) {
        const double weight = smax(0., w[k]);
        value += weight * (double)density[V.tetv[4 * t + k]];
        total += weight;
    }
    return value;
}

// One dual face's circulation (SimplicialFluidSolver3D.cpp:444-449 and
// DualMesh::circulation, SimplicialDual3D.cpp:161-166): each traced dual edge
// by the trapezoid rule, summed around the loop in order.
__host__ __device__ __forceinline__ double loop_circulation(const int* loop_edges, const int* loop_signs, int j0, int j1,
                                                            const int* dual_edges, const double* traced,
                                                            const double* velocity) {
    double omega = 0;
    for (int j = j0; j < j1; ++j) {
        const int e = loop_edges[j];
        const int a = dual_edges[2 * e], b = dual_edges[2 * e + 1];
        const double integral =
            .5 * dot(add(ld3(velocity, a), ld3(velocity, b)), sub(ld3(traced, b), ld3(traced, a)));
        omega += loop_signs[j] * integral;
    }
    return omega;
}

// ---------------------------------------------------------------------------
// One traced point per call, shared by the kernels (one thread each) and a
// CPU build of them. `nearest(p)` returns the nearest primal vertex as
// nearer() orders them.

// Box trace of p (SimplicialFluidSolver3D.cpp:416-421): midpoint rule with
// both points clamped to the box; `active` false skips the interpolations
// (a lane past the end that still takes part in wave-uniform scans).
template <class Nearest>
__host__ __device__ __forceinline__ P3 box_trace(const View& V, Nearest&& nearest, P3 p, const double* velocity,
                                                 double dt, bool active, int& error) {
    int site = nearest(p);
    P3 v = active ? warren(V, site, p, velocity, error) : P3{0, 0, 0};
    const P3 midpoint = clamp_box(sub(p, mul(v, dt * .5)), V.size);
    site = nearest(midpoint);
    v = active ? warren(V, site, midpoint, velocity, error) : P3{0, 0, 0};
    return clamp_box(sub(p, mul(v, dt)), V.size);
}

// Box smoke at primal vertex i (SimplicialFluidSolver3D.cpp:600, sampleDensity :500-511).
template <class Nearest>
__host__ __device__ __forceinline__ float box_density_point(const View& V, Nearest&& nearest, int i,
                                                           const double* velocity, const float* density, double dt,
                                                           double decay, bool active, int& error) {
    const P3 q = box_trace(V, nearest, active ? ld3(V.vert, i) : P3{0, 0, 0}, velocity, dt, active, error);
    if (!active) return 0.f;
    const int t = locate(V, q, -1);
    double sample = 0;
    if (t >= 0) {
        double total;
        const double value = tet_density(V, t, q, density, total);
        sample = value / total;
    }
    return (float)(sample * decay);
}

// General domain, dual vertex i (SimplicialFluidSolver3D.cpp:439-444).
__host__ __device__ inline void general_trace_point(const View& V, int i, const double* velocity,
                                                    const double* samples, double dt, P3& q, P3& v, int& error) {
    const P3 p = ld3(V.dual_pos, i);
    const int hint = V.dual_hint[i];
    const P3 midpoint = clip_segment(V, p, sub(p, mul(ld3(velocity, i), dt * .5)), hint);
    const P3 middle_velocity = subdivision_value(V, midpoint, hint, samples, error);
    q = clip_segment(V, p, sub(p, mul(middle_velocity, dt)), hint);
    v = subdivision_value(V, q, hint, samples, error);
}

// General domain, primal vertex i (SimplicialFluidSolver3D.cpp:601-609).
__host__ __device__ inline float general_density_point(const View& V, int i, const double* samples,
                                                       const float* density, double dt, double decay, int& error) {
    const P3 p = ld3(V.vert, i);
    const int hint = V.primal_hint[i];
    const P3 middle = clip_segment(V, p, sub(p, mul(ld3(samples, i), dt * .5)), hint);
    const P3 velocity = subdivision_value(V, middle, hint, samples, error);
    const P3 q = clip_segment(V, p, sub(p, mul(velocity, dt)), hint);
    const int t = locate(V, q, hint);
    double value = 0, total = 0;
    if (t >= 0) value = tet_density(V, t, q, density, total);
    return total > 0 ? (float)(value / total * decay) : 0.f;
}

// ---------------------------------------------------------------------------
// The potential operator of SimplicialFluidSolver3D::applyLaplacian
// (SimplicialFluidSolver3D.cpp:106-129), gathered so that each sum runs in
// the reference's order: per face its three edges and signs; per vertex its
// edges in ascending order (item e adds star1 x, ~e subtracts it); per edge
// its faces in ascending order (item f adds c[f], ~f subtracts it). The
// IC(0) factor as lower rows (ascending columns) and as upper rows
// (transposed, descending rows), each with a level schedule.
struct CgOperator {
    int ne, nf, nv;
    const int* face_edges;
    const double* face_signs;
    const double* star2;
    const int* vertex_off;
    const int* vertex_items;
    const double* star1;
    const double* star0;
    const unsigned char* vertex_boundary;
    const int* edge_off;
    const int* edge_items;
    const int* edge_vertices;
    const unsigned char* edge_boundary;
    const double* diagonal;
    int forward_levels, backward_levels;  // 0 without an IC factor
    const int* forward_off;
    const int* forward_rows;
    const int* lower_off;
    const int* lower_columns;
    const double* lower_values;
    const int* backward_off;
    const int* backward_rows;
    const int* upper_off;
    const int* upper_rows;
    const double* upper_values;
    const double* ic_diagonal;
};

// The operator rows read u through an accessor u(e) (a plain array, or a
// vector formed on the fly such as z + beta p).
//
// d1 then star2 for face f of u with its boundary entries read as zero
// (Mesh::d1, SimplicialMesh3D.cpp:154-157; curl's star2, :98-105).
template <class U>
__host__ __device__ __forceinline__ double face_value(const CgOperator& A, const U& u, int f) {
    double s = 0;
    for (int k = 0; k < 3; ++k) {
        const int e = A.face_edges[3 * f + k];
        const double x = A.edge_boundary[e] ? 0. : u(e);
        s += A.face_signs[3 * f + k] * x;
    }
    return s * A.star2[f];
}

// The divergence at vertex v: star1 x over its edges in ascending edge
// order, divided by star0, zero on the boundary (:120-128). The gather runs
// in chunks of four: the chunk's loads are issued together, the sum still
// takes the terms one by one in order.
template <class U>
__host__ __device__ __forceinline__ double vertex_divergence(const CgOperator& A, const U& u, int v) {
    double acc = 0;
    const int j0 = A.vertex_off[v], j1 = A.vertex_off[v + 1];
    for (int j = j0; j < j1; j += 4) {
        int item[4];
        double value[4];
        for (int k = 0; k < 4; ++k) item[k] = A.vertex_items[j + k < j1 ? j + k : j0];
        for (int k = 0; k < 4; ++k) {
            const int e = item[k] >= 0 ? item[k] : ~item[k];
            value[k] = A.star1[e] * (A.edge_boundary[e] ? 0. : u(e));
        }
        for (int k = 0; k < 4; ++k)
            if (j + k < j1) {
                if (item[k] >= 0) acc += value[k];
                else acc -= value[k];
            }
    }
    return A.vertex_boundary[v] ? 0 : acc / A.star0[v];
}

// vertex_divergence on kVertexLanes lanes at once (device only): lane k of
// the group loads the group's k-th edge term, then every lane folds the terms
// in ascending edge order through shuffles - the same sum, in the same order,
// with the loads of a chunk in flight together instead of one chain per
// vertex. The group's lanes are consecutive lanes of one wave and call this
// together.
constexpr int kVertexLanes = 16;
template <class U>
__device__ __forceinline__ double vertex_divergence_lanes(const CgOperator& A, const U& u, int v, int k) {
    double acc = 0;
    const int j0 = A.vertex_off[v], j1 = A.vertex_off[v + 1];
    const int base = (int)(threadIdx.x & 63) & ~(kVertexLanes - 1);
    for (int j = j0; j < j1; j += kVertexLanes) {
        int item = 0;
        double value = 0;
        if (j + k < j1) {
            item = A.vertex_items[j + k];
            const int e = item >= 0 ? item : ~item;
            value = A.star1[e] * (A.edge_boundary[e] ? 0. : u(e));
        }
        const int count = j1 - j < kVertexLanes ? j1 - j : kVertexLanes;
        for (int m = 0; m < count; ++m) {
            const double t = __shfl(value, base + m);
            if (__shfl(item, base + m) >= 0) acc += t;
            else acc -= t;
        }
    }
    return A.vertex_boundary[v] ? 0 : acc / A.star0[v];
}

// Row e of applyLaplacian for an edge whose own entry is ue: d1^T of the
// face values in ascending face order (Mesh::d1Transpose,
// SimplicialMesh3D.cpp:162-165), plus star1 d0 of the divergence; boundary
// rows return ue. The faces are gathered in chunks of kEdgeChunk (an edge of
// a tetrahedral mesh lies in about five faces, so most take one chunk): the
// chunk's loads are issued together, the sum still runs in order.
constexpr int kEdgeChunk = 8;
__host__ __device__ __forceinline__ double edge_value(const CgOperator& A, double ue, const double* c,
                                                      const double* div, int e) {
    double acc = 0;
    const int j0 = A.edge_off[e], j1 = A.edge_off[e + 1];
    for (int j = j0; j < j1; j += kEdgeChunk) {
        int item[kEdgeChunk];
        double value[kEdgeChunk];
        for (int k = 0; k < kEdgeChunk; ++k) item[k] = A.edge_items[j + k < j1 ? j + k : j0];
        for (int k = 0; k < kEdgeChunk; ++k) value[k] = c[item[k] >= 0 ? item[k] : ~item[k]];
        for (int k = 0; k < kEdgeChunk; ++k)
            if (j + k < j1) acc += (item[k] >= 0 ? 1. : -1.) * value[k];
    }
    acc += A.star1[e] * (div[A.edge_vertices[2 * e + 1]] - div[A.edge_vertices[2 * e]]);
    return A.edge_boundary[e] ? ue : acc;
}

// An array as an accessor.
struct Entries {
    const double* a;
    __host__ __device__ __forceinline__ double operator()(int e) const { return a[e]; }
};

// Row i of the IC(0) forward solve and row j of the backward solve
// (SimplicialFluidSolver3D.cpp:166-173), each in the reference's order.
__host__ __device__ __forceinline__ double ic_forward_row(const CgOperator& A, const double* r, const double* z, int i) {
    double zi = r[i];
    for (int j = A.lower_off[i]; j < A.lower_off[i + 1]; ++j) zi -= A.lower_values[j] * z[A.lower_columns[j]];
    return zi / A.ic_diagonal[i];
}
__host__ __device__ __forceinline__ double ic_backward_row(const CgOperator& A, const double* z, int j) {
    double zj = z[j];
    for (int m = A.upper_off[j]; m < A.upper_off[j + 1]; ++m) zj -= A.upper_values[m] * z[A.upper_rows[m]];
    return zj / A.ic_diagonal[j];
}

}  // namespace vsx
