// vsx_kernels.hip - the simplicial 3D backend's kernels on gfx950.
//
// Advection: a box trace looks up the nearest primal vertex three times by
// the reference's brute-force scan, split over G lanes per traced point with
// the vertices staged in LDS tiles; each lane's work is a distance, a compare
// and two selects per vertex. General domains locate, clip and interpolate
// through the subdivision, one lane per point.
// Recovery: the Laplacian is applied matrix-free with every sum in the
// reference's order. In the reference recovery order one workgroup of 1024
// lanes runs the whole CG with index-order inner products (bit-identical to
// the CPU); in the fast order one cooperative grid over the CUs runs it with
// fixed-order tree reductions.
#include <hip/hip_cooperative_groups.h>

#include "vsx_kernels.hpp"

namespace vsx {

namespace {

// Nearest primal vertex to p on G lanes (SimplicialDual3D.cpp:155-156: a
// strict-less scan over every vertex, the first nearest in index order wins).
// The workgroup stages the vertices in LDS tiles of kTraceBlock; lane k of a
// point's G consecutive lanes scans the tile entries k, k + G, ... in order
// (its own first nearest), then the G lanes merge (distance, index) pairs:
// the smaller distance wins, the smaller index at equal distances - the
// first nearest overall. Every lane of the workgroup calls this together.
template <int G>
__device__ __forceinline__ int nearest_site(const double* __restrict__ vert, int nv, P3 p, double* tile) {
    double best = 1.7976931348623157e308;  // std::numeric_limits<double>::max()
    int site = 0;
    const int k = threadIdx.x % G;
    for (int base = 0; base < nv; base += kTraceBlock) {
        __syncthreads();  // the previous tile's reads are done
        const int i = base + (int)threadIdx.x;
        if (i < nv) {
            tile[3 * threadIdx.x] = vert[3 * i];
            tile[3 * threadIdx.x + 1] = vert[3 * i + 1];
            tile[3 * threadIdx.x + 2] = vert[3 * i + 2];
        }
        __syncthreads();
        const int count = nv - base < kTraceBlock ? nv - base : kTraceBlock;
        for (int j = k; j < count; j += G) nearer(p, ld3(tile, j), base + j, best, site);
    }
    for (int o = G / 2; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o);
        const int os = __shfl_xor(site, o);
        if (ob < best || (ob == best && os < site)) {
            best = ob;
            site = os;
        }
    }
    return site;
}

}  // namespace

// Box domains, G lanes per traced point. DENSITY false: dual vertex i ->
// traced point and velocity there (SimplicialFluidSolver3D.cpp:434-437).
// DENSITY true: primal vertex i -> float(sample(trace(p)) * decay)
// (SimplicialFluidSolver3D.cpp:600). Every lane runs the scans (workgroup
// barriers inside); points past the end skip only their interpolations and
// stores. The point's lanes interpolate alike; its first lane stores.
template <bool DENSITY, int G>
__global__ void __launch_bounds__(kTraceBlock) k_box_trace(View V, const double* __restrict__ vert,
                                                           const double* __restrict__ velocity, double dt, double decay,
                                                           const float* __restrict__ density, double* __restrict__ traced,
                                                           double* __restrict__ traced_velocity,
                                                           float* __restrict__ density_out, int* __restrict__ error) {
    __shared__ double tile[3 * kTraceBlock];
    const int count = DENSITY ? V.nv : V.nd;
    const int i = (blockIdx.x * kTraceBlock + (int)threadIdx.x) / G;
    const bool active = i < count;
    const bool writer = active && threadIdx.x % G == 0;
    int err = 0;
    auto nearest = [&](P3 p) { return nearest_site<G>(vert, V.nv, p, tile); };
    if (DENSITY) {
        const float value = box_density_point(V, nearest, i, velocity, density, dt, decay, active, err);
        if (writer) density_out[i] = value;
    } else {
        const P3 q = box_trace(V, nearest, active ? ld3(V.dual_pos, i) : P3{0, 0, 0}, velocity, dt, active, err);
        const int site = nearest(q);
        if (active) {
            const P3 v = warren(V, site, q, velocity, err);
            if (writer) {
                st3(traced, i, q);
                st3(traced_velocity, i, v);
            }
        }
    }
    if (err) atomicOr(error, err);
}

// General domains, dual vertex i.
__global__ void __launch_bounds__(kGeneralBlock) k_general_trace(View V, const double* __restrict__ velocity,
                                                               const double* __restrict__ samples, double dt,
                                                               double* __restrict__ traced,
                                                               double* __restrict__ traced_velocity,
                                                               int* __restrict__ error) {
    const int i = blockIdx.x * kGeneralBlock + threadIdx.x;
    if (i >= V.nd) return;
    int err = 0;
    P3 q, v;
    general_trace_point(V, i, velocity, samples, dt, q, v, err);
    st3(traced, i, q);
    st3(traced_velocity, i, v);
    if (err) atomicOr(error, err);
}

// General domains, primal vertex i.
__global__ void __launch_bounds__(kGeneralBlock) k_general_density(View V, const double* __restrict__ samples,
                                                                 const float* __restrict__ density, double dt,
                                                                 double decay, float* __restrict__ out,
                                                                 int* __restrict__ error) {
    const int i = blockIdx.x * kGeneralBlock + threadIdx.x;
    if (i >= V.nv) return;
    int err = 0;
    out[i] = general_density_point(V, i, samples, density, dt, decay, err);
    if (err) atomicOr(error, err);
}

// One primal edge's dual-face circulation.
__global__ void __launch_bounds__(256) k_loops(int ne, const int* __restrict__ loop_off,
                                               const int* __restrict__ loop_edges, const int* __restrict__ loop_signs,
                                               const int* __restrict__ dual_edges, const double* __restrict__ traced,
                                               const double* __restrict__ velocity, double* __restrict__ omega) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ne) return;
    omega[e] = loop_circulation(loop_edges, loop_signs, loop_off[e], loop_off[e + 1], dual_edges, traced, velocity);
}

// ---------------------------------------------------------------------------
// Recovery in the reference order: SimplicialFluidSolver3D::solve
// (SimplicialFluidSolver3D.cpp:135-217) in one workgroup, every inner product
// summed in index order by one lane, the IC(0) factor where there is one.

namespace {

// max(s, |x|) as `std::max(s, std::abs(x))` combines: NaN never replaces s.
__device__ __forceinline__ double absmax(double s, double x) {
    const double a = __builtin_fabs(x);
    return (s < a) ? a : s;
}

__device__ __forceinline__ double block_max(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v = absmax(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0;
    for (int w = 0; w < kCgBlock / 64; ++w) s = absmax(s, red[w]);
    return s;
}

// y = applyLaplacian(u) (SimplicialFluidSolver3D.cpp:106-129): boundary
// entries of u read as zero, d1, star2, d1^T gathered per edge in ascending
// face order, the vertex divergence gathered in ascending edge order,
// star1 d0; boundary rows return u.
__device__ void apply_laplacian(const CgOperator& A, const double* __restrict__ u, double* __restrict__ y,
                                double* __restrict__ c, double* __restrict__ div) {
    const Entries ue{u};
    for (int f = threadIdx.x; f < A.nf; f += kCgBlock) c[f] = face_value(A, ue, f);
    for (int slot = threadIdx.x; slot < kVertexLanes * A.nv; slot += kCgBlock) {
        const int v = slot / kVertexLanes, k = slot % kVertexLanes;
        const double d = vertex_divergence_lanes(A, ue, v, k);
        if (k == 0) div[v] = d;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < A.ne; e += kCgBlock) y[e] = edge_value(A, u[e], c, div, e);
    __syncthreads();
}

// inner(a, b) as std::inner_product sums it: a[i] * b[i] from 0 in index
// order, in one lane (the products are formed in parallel first).
__device__ double inner(const CgOperator& A, const double* __restrict__ a, const double* __restrict__ b,
                        double* __restrict__ prod, double* red) {
    for (int e = threadIdx.x; e < A.ne; e += kCgBlock) prod[e] = a[e] * b[e];
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int e = 0; e < A.ne; ++e) s = s + prod[e];
        red[kCgBlock / 64] = s;
    }
    __syncthreads();
    const double s = red[kCgBlock / 64];
    __syncthreads();
    return s;
}

__device__ double norm(const CgOperator& A, const double* __restrict__ a, double* red) {
    double s = 0;
    for (int e = threadIdx.x; e < A.ne; e += kCgBlock) s = absmax(s, a[e]);
    return block_max(s, red);
}

// z = M^-1 r: the Jacobi diagonal, or the IC(0) factor's two triangular
// solves (SimplicialFluidSolver3D.cpp:163-177) level by level, each row in
// the reference's order.
template <bool IC>
__device__ void precondition(const CgOperator& A, const double* __restrict__ r, double* __restrict__ z) {
    if (!IC) {
        for (int e = threadIdx.x; e < A.ne; e += kCgBlock) z[e] = r[e] / A.diagonal[e];
        __syncthreads();
        return;
    }
    for (int level = 0; level < A.forward_levels; ++level) {
        for (int k = A.forward_off[level] + threadIdx.x; k < A.forward_off[level + 1]; k += kCgBlock) {
            const int i = A.forward_rows[k];
            z[i] = ic_forward_row(A, r, z, i);
        }
        __syncthreads();
    }
    for (int level = 0; level < A.backward_levels; ++level) {
        for (int k = A.backward_off[level] + threadIdx.x; k < A.backward_off[level + 1]; k += kCgBlock) {
            const int j = A.backward_rows[k];
            z[j] = ic_backward_row(A, z, j);
        }
        __syncthreads();
    }
}

}  // namespace

template <bool IC>
__global__ void __launch_bounds__(kCgBlock) k_cg(CgOperator A, CgVectors W, CgControl control,
                                                 CgStatus* __restrict__ status) {
    __shared__ double red[kCgBlock / 64 + 1];
    __shared__ int flag;
    const int ne = A.ne;
    double *x = W.x, *r = W.r, *z = W.z, *p = W.p, *ap = W.ap, *rhs = W.rhs;
    if (threadIdx.x == 0) flag = 0;
    __syncthreads();
    // x0 and r0 (:146-151).
    if (control.warm_start) {
        for (int e = threadIdx.x; e < ne; e += kCgBlock) x[e] = W.phi[e];
        __syncthreads();
        apply_laplacian(A, x, ap, W.c, W.div);
        for (int e = threadIdx.x; e < ne; e += kCgBlock) r[e] = rhs[e] - ap[e];
    } else {
        for (int e = threadIdx.x; e < ne; e += kCgBlock) {
            x[e] = 0;
            r[e] = rhs[e];
        }
    }
    // Finite check and boundary rows (:152-160).
    int bad = 0;
    for (int e = threadIdx.x; e < ne; e += kCgBlock) {
        if (!(r[e] - r[e] == 0)) bad = 1;
        if (A.edge_boundary[e]) {
            r[e] = 0;
            rhs[e] = 0;
        }
    }
    if (bad) atomicOr(&flag, 1);
    __syncthreads();
    if (flag) {
        if (threadIdx.x == 0) status->error = kErrNonFiniteRhs;
        return;
    }
    const double initial = norm(A, r, red);
    const double scale = control.warm_start ? norm(A, rhs, red) : initial;
    const double target = smax(control.absolute_tolerance, control.relative_tolerance * scale);
    double final_residual = initial;
    int iterations = 0;
    if (initial <= target) {
        if (threadIdx.x == 0) {
            status->error = 0;
            status->iterations = 0;
            status->converged = 1;
            status->initial = initial;
            status->final = initial;
        }
        return;
    }
    precondition<IC>(A, r, z);
    for (int e = threadIdx.x; e < ne; e += kCgBlock) p[e] = z[e];
    __syncthreads();
    double rz = inner(A, r, z, W.prod, red);
    for (int it = 0; it < control.max_iterations; ++it) {
        apply_laplacian(A, p, ap, W.c, W.div);
        const double pap = inner(A, p, ap, W.prod, red);
        if (!(pap - pap == 0) || pap <= 0) {
            if (threadIdx.x == 0) status->error = kErrDefinite;
            return;
        }
        const double alpha = rz / pap;
        for (int e = threadIdx.x; e < ne; e += kCgBlock) {
            x[e] += alpha * p[e];
            r[e] -= alpha * ap[e];
        }
        __syncthreads();
        iterations = it + 1;
        final_residual = norm(A, r, red);
        if (final_residual <= target) break;
        precondition<IC>(A, r, z);
        const double next = inner(A, r, z, W.prod, red), beta = next / rz;
        for (int e = threadIdx.x; e < ne; e += kCgBlock) p[e] = z[e] + beta * p[e];
        __syncthreads();
        rz = next;
    }
    // The reported residual is recomputed from x (:207-213).
    apply_laplacian(A, x, ap, W.c, W.div);
    for (int e = threadIdx.x; e < ne; e += kCgBlock) r[e] = A.edge_boundary[e] ? 0 : rhs[e] - ap[e];
    __syncthreads();
    final_residual = norm(A, r, red);
    if (threadIdx.x == 0) {
        status->error = 0;
        status->iterations = iterations;
        status->converged = final_residual <= target ? 1 : 0;
        status->initial = initial;
        status->final = final_residual;
    }
}

// ---------------------------------------------------------------------------
// Recovery over the whole GPU (fast order): the same CG with Jacobi
// preconditioning as one cooperative kernel. Every grid-stride loop gives
// each lane about one face, vertex or edge; a reduction writes one partial
// per workgroup, and after the grid barrier every workgroup sums the
// partials in workgroup order, so all workgroups hold the same scalars and
// take the same branches (every exit below is grid-uniform).

namespace {

namespace cgr = cooperative_groups;

// This workgroup's sum (or NaN-ignoring max) of v, in a fixed order.
template <bool MAX>
__device__ __forceinline__ double group_reduce(double v, double* lds) {
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o);
        v = MAX ? absmax(v, w) : v + w;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0;
    for (int w = 0; w < kGridBlock / 64; ++w) s = MAX ? absmax(s, lds[w]) : s + lds[w];
    return s;
}

// The grid's total from one partial per workgroup (after a grid barrier):
// lane l of each wave folds partials l, l+64, ... then the wave's tree.
template <bool MAX>
__device__ __forceinline__ double grid_total(const double* partials, int groups) {
    const int lane = threadIdx.x & 63;
    double v = 0;
    for (int g = lane; g < groups; g += 64) v = MAX ? absmax(v, partials[g]) : v + partials[g];
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o);
        v = MAX ? absmax(v, w) : v + w;
    }
    return v;
}

}  // namespace

// Structure: the setup forms x0, r0, z0 and p0 = z0; each iteration then
// takes three grid barriers - after the face and vertex values of p, after
// Ap and the partials of p.Ap, after the x / r / z update and the partials of
// |r| and r.z. The direction update p = z + beta p_old of
// SimplicialFluidSolver3D.cpp:203-204 is not a pass of its own: the next
// iteration forms each entry where it reads it (the same operations, so the
// same bits), and the edge that owns it stores it into the other of two p
// buffers.
__global__ void __launch_bounds__(kGridBlock) k_cg_grid(CgOperator A, CgVectors W, CgControl control, GridPartials G,
                                                        CgStatus* __restrict__ status) {
    __shared__ double lds[kGridBlock / 64];
    cgr::grid_group grid = cgr::this_grid();
    const int groups = gridDim.x;
    const int first = blockIdx.x * kGridBlock + threadIdx.x, stride = groups * kGridBlock;
    const bool leader = threadIdx.x == 0;
    const int ne = A.ne;
    double *x = W.x, *r = W.r, *z = W.z, *ap = W.ap, *rhs = W.rhs;
    double* pbuf[2] = {W.p, W.prod};
    // y = applyLaplacian(u) for an accessor u (one barrier, between the face
    // and vertex values and the edge rows).
    auto apply = [&](const auto& u, double* y) {
        for (int f = first; f < A.nf; f += stride) W.c[f] = face_value(A, u, f);
        for (int slot = first; slot < kVertexLanes * A.nv; slot += stride) {
            const int v = slot / kVertexLanes, k = slot % kVertexLanes;
            const double d = vertex_divergence_lanes(A, u, v, k);
            if (k == 0) W.div[v] = d;
        }
        grid.sync();
        for (int e = first; e < ne; e += stride) y[e] = edge_value(A, u(e), W.c, W.div, e);
    };
    // x0, r0 and the finite check (SimplicialFluidSolver3D.cpp:146-160).
    if (control.warm_start) {
        for (int e = first; e < ne; e += stride) x[e] = W.phi[e];
        grid.sync();
        apply(Entries{x}, ap);
        for (int e = first; e < ne; e += stride) r[e] = rhs[e] - ap[e];
    } else {
        for (int e = first; e < ne; e += stride) {
            x[e] = 0;
            r[e] = rhs[e];
        }
    }
    double bad = 0, rmax = 0, bmax = 0, rz = 0;
    for (int e = first; e < ne; e += stride) {
        if (!(r[e] - r[e] == 0)) bad = 1;
        if (A.edge_boundary[e]) r[e] = rhs[e] = 0;
        rmax = absmax(rmax, r[e]);
        bmax = absmax(bmax, rhs[e]);
        // z0 = M^-1 r0 and p0 = z0, used only if the solve starts.
        const double ze = r[e] / A.diagonal[e];
        z[e] = ze;
        pbuf[0][e] = ze;
        rz += r[e] * ze;
    }
    bad = group_reduce<true>(bad, lds);
    rmax = group_reduce<true>(rmax, lds);
    bmax = group_reduce<true>(bmax, lds);
    rz = group_reduce<false>(rz, lds);
    if (leader) {
        G.a[blockIdx.x] = bad;
        G.b[blockIdx.x] = rmax;
        G.c[blockIdx.x] = bmax;
        G.d[blockIdx.x] = rz;
    }
    grid.sync();
    if (grid_total<true>(G.a, groups) != 0) {
        if (blockIdx.x == 0 && leader) status->error = kErrNonFiniteRhs;
        return;
    }
    const double initial = grid_total<true>(G.b, groups);
    const double scale = control.warm_start ? grid_total<true>(G.c, groups) : initial;
    const double target = smax(control.absolute_tolerance, control.relative_tolerance * scale);
    if (initial <= target) {
        if (blockIdx.x == 0 && leader) {
            status->error = 0;
            status->iterations = 0;
            status->converged = 1;
            status->initial = initial;
            status->final = initial;
        }
        return;
    }
    rz = grid_total<false>(G.d, groups);
    int iterations = 0;
    double beta = 0;  // p of iteration 0 is z + 0 p0 = z0 exactly (p0 = z0)
    for (int it = 0; it < control.max_iterations; ++it) {
        const double* p_old = pbuf[it & 1];
        double* p_new = pbuf[(it + 1) & 1];
        const auto pe = [&](int e) { return z[e] + beta * p_old[e]; };
        apply(pe, ap);
        double pap = 0;
        for (int e = first; e < ne; e += stride) {
            const double p = pe(e);
            p_new[e] = p;
            pap += p * ap[e];
        }
        pap = group_reduce<false>(pap, lds);
        if (leader) G.a[blockIdx.x] = pap;
        grid.sync();
        pap = grid_total<false>(G.a, groups);
        if (!(pap - pap == 0) || pap <= 0) {
            if (blockIdx.x == 0 && leader) status->error = kErrDefinite;
            return;
        }
        const double alpha = rz / pap;
        double residual = 0, next = 0;
        for (int e = first; e < ne; e += stride) {
            x[e] += alpha * p_new[e];
            const double re = r[e] - alpha * ap[e];
            r[e] = re;
            residual = absmax(residual, re);
            const double ze = re / A.diagonal[e];
            z[e] = ze;
            next += re * ze;
        }
        residual = group_reduce<true>(residual, lds);
        next = group_reduce<false>(next, lds);
        if (leader) {
            G.b[blockIdx.x] = residual;
            G.c[blockIdx.x] = next;
        }
        grid.sync();
        iterations = it + 1;
        if (grid_total<true>(G.b, groups) <= target) break;
        next = grid_total<false>(G.c, groups);
        beta = next / rz;
        rz = next;
    }
    // The reported residual is recomputed from x (:207-213).
    apply(Entries{x}, ap);
    double residual = 0;
    for (int e = first; e < ne; e += stride) residual = absmax(residual, A.edge_boundary[e] ? 0. : rhs[e] - ap[e]);
    residual = group_reduce<true>(residual, lds);
    if (leader) G.d[blockIdx.x] = residual;
    grid.sync();
    const double final_residual = grid_total<true>(G.d, groups);
    if (blockIdx.x == 0 && leader) {
        status->error = 0;
        status->iterations = iterations;
        status->converged = final_residual <= target ? 1 : 0;
        status->initial = initial;
        status->final = final_residual;
    }
}

// ---------------------------------------------------------------------------
// Launchers.

static inline int blocks_for(int count, int block) { return (count + block - 1) / block; }

// Lanes per traced point: enough lanes in all to give every SIMD a few waves
// (small meshes have few points), at most 32.
template <bool DENSITY>
static void launch_box_trace(const View& V, const double* velocity, double dt, double decay, const float* density,
                             double* traced, double* traced_velocity, float* out, int* error, hipStream_t stream) {
    const int count = DENSITY ? V.nv : V.nd;
    const long target = 1L << 18;
    int g = 1;
    while (g < 32 && (long)count * g < target) g *= 2;
#define VSX_BOX_TRACE(G)                                                                                        \
    hipLaunchKernelGGL((k_box_trace<DENSITY, G>), dim3(blocks_for(count * G, kTraceBlock)), dim3(kTraceBlock), 0, \
                       stream, V, V.vert, velocity, dt, decay, density, traced, traced_velocity, out, error)
    switch (g) {
    case 1: VSX_BOX_TRACE(1); break;
    case 2: VSX_BOX_TRACE(2); break;
    case 4: VSX_BOX_TRACE(4); break;
    case 8: VSX_BOX_TRACE(8); break;
    case 16: VSX_BOX_TRACE(16); break;
    default: VSX_BOX_TRACE(32); break;
    }
#undef VSX_BOX_TRACE
}

void launch_circulation(const View& V, const double* velocity, const double* samples, double dt, double* traced,
                        double* traced_velocity, const int* loop_off, const int* loop_edges, const int* loop_signs,
                        const int* dual_edges, double* omega, int* error, hipStream_t stream) {
    if (V.box)
        launch_box_trace<false>(V, velocity, dt, 1.0, nullptr, traced, traced_velocity, nullptr, error, stream);
    else
        hipLaunchKernelGGL(k_general_trace, dim3(blocks_for(V.nd, kGeneralBlock)), dim3(kGeneralBlock), 0, stream, V,
                           velocity, samples, dt, traced, traced_velocity, error);
    hipLaunchKernelGGL(k_loops, dim3(blocks_for(V.ne, 256)), dim3(256), 0, stream, V.ne, loop_off, loop_edges,
                       loop_signs, dual_edges, traced, traced_velocity, omega);
}

void launch_density(const View& V, const double* velocity, const double* samples, const float* density, double dt,
                    double decay, float* out, int* error, hipStream_t stream) {
    if (V.box)
        launch_box_trace<true>(V, velocity, dt, decay, density, nullptr, nullptr, out, error, stream);
    else
        hipLaunchKernelGGL(k_general_density, dim3(blocks_for(V.nv, kGeneralBlock)), dim3(kGeneralBlock), 0, stream, V,
                           samples, density, dt, decay, out, error);
}

void launch_cg(const CgOperator& A, const CgVectors& W, const CgControl& control, CgStatus* status,
               hipStream_t stream) {
    if (A.forward_levels > 0)
        hipLaunchKernelGGL(k_cg<true>, dim3(1), dim3(kCgBlock), 0, stream, A, W, control, status);
    else
        hipLaunchKernelGGL(k_cg<false>, dim3(1), dim3(kCgBlock), 0, stream, A, W, control, status);
}

int cg_grid_capacity(int device) {
    int per_cu = 0;
    hipDeviceProp_t prop;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_cg_grid), kGridBlock,
                                                     0) != hipSuccess ||
        hipGetDeviceProperties(&prop, device) != hipSuccess)
        return 0;
    // One workgroup per CU at most: enough lanes for the meshes this serves,
    // and a cheaper grid barrier than a fuller grid.
    return std::min(per_cu, 1) * prop.multiProcessorCount;
}

int cg_grid_groups(const CgOperator& A, int max_groups) {
    const int items = std::max(A.nf, std::max(A.ne, A.nv));
    return std::max(1, std::min(max_groups, (items + kGridBlock - 1) / kGridBlock));
}

hipError_t launch_cg_grid(const CgOperator& A, const CgVectors& W, const CgControl& control, const GridPartials& G,
                          int groups, CgStatus* status, hipStream_t stream) {
    CgOperator a = A;
    CgVectors w = W;
    CgControl c = control;
    GridPartials g = G;
    void* args[] = {&a, &w, &c, &g, &status};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_cg_grid), dim3(groups), dim3(kGridBlock), args,
                                      0, stream);
}

}  // namespace vsx
