// Face arrays of a finite-volume mesh built on the device from its cells:
// mesh/mesh_reader.py:141-242 (build_fvmesh_from_arrays, 2D edges) and
// :245-323 (_build_fvmesh_3d, per-cell face lists).
//
// The reference walks the cells in order and numbers a face the first time
// it meets its node set (a dict keyed by the sorted node ids); the cell that
// first meets it owns it, and a later cell that meets it becomes its
// neighbour (the last such cell, since each encounter overwrites).  Here
// every (cell, local face) entry gets its sorted-node key, a stable radix
// sort brings equal keys together in entry order, and a scan over the
// "first entry of its key" flags numbers the faces in first-encounter
// order.  Geometry (Newell normal and area, centres, divergence-theorem
// volume, 2D shoelace) follows the vectorised host builder's evaluation
// order operation by operation, with floating-point contraction off, so the
// arrays equal the host's bit for bit.  Boundary patches: a boundary list
// entry names the faces with its key, a later list entry overriding an
// earlier one (mesh_reader.py:218-230, 306-315).
#pragma clang fp contract(off)

#include <hipcub/hipcub.hpp>

#include <cstring>
#include <memory>

#include "common.hpp"

using namespace kcfd;

namespace {

constexpr unsigned kNoNode = 0x7fffffffu;   // padding in a key: after any node id
constexpr int kMbBlock = 256;

inline int mb_grid(long long n) { return (int)std::max(1LL, (n + kMbBlock - 1) / kMbBlock); }

struct Key {
    unsigned long long hi, lo;
};

__device__ __forceinline__ void sort4(unsigned &a, unsigned &b, unsigned &c, unsigned &d) {
    unsigned t;
#define KCFD_CSWAP(x, y) if (x > y) { t = x; x = y; y = t; }
    KCFD_CSWAP(a, b) KCFD_CSWAP(c, d) KCFD_CSWAP(a, c) KCFD_CSWAP(b, d) KCFD_CSWAP(b, c)
#undef KCFD_CSWAP
}

__device__ __forceinline__ Key make_key(const int *nd, int w) {
    unsigned k[4] = {kNoNode, kNoNode, kNoNode, kNoNode};
    for (int i = 0; i < w; ++i)
        if (nd[i] >= 0) k[i] = (unsigned)nd[i];
    sort4(k[0], k[1], k[2], k[3]);
    Key r;
    r.hi = ((unsigned long long)k[0] << 32) | k[1];
    r.lo = ((unsigned long long)k[2] << 32) | k[3];
    return r;
}

// the nodes of entry e: 2D, edge k of cell c = e / npc (nodes k, k+1 mod
// npc); 3D, row e of the per-cell face lists
__device__ __forceinline__ void entry_nodes(int e, int ndim, const int *cells, int npc,
                                            const int *faces, int fw, int *out) {
    if (ndim == 2) {
        const int c = e / npc, k = e - c * npc;
        out[0] = cells[(size_t)c * npc + k];
        out[1] = cells[(size_t)c * npc + (k + 1 == npc ? 0 : k + 1)];
        out[2] = out[3] = -1;
    } else {
        for (int i = 0; i < 4; ++i) out[i] = i < fw ? faces[(size_t)e * fw + i] : -1;
    }
}

__global__ void k_mb_keys(int E, int ndim, const int *cells, int npc, const int *faces, int fw,
                          unsigned long long *hi, unsigned long long *lo, int *idx) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    int nd[4];
    entry_nodes(e, ndim, cells, npc, faces, fw, nd);
    const Key k = make_key(nd, 4);
    hi[e] = k.hi;
    lo[e] = k.lo;
    idx[e] = e;
}

// cell of each entry (3D: from the per-cell offsets)
__global__ void k_mb_entry_cell(int n, const int *ptr, int *ent_cell) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int e = ptr[c]; e < ptr[c + 1]; ++e) ent_cell[e] = c;
}

__global__ void k_mb_gather(int E, const int *idx, const unsigned long long *src, unsigned long long *dst) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < E) dst[s] = src[idx[s]];
}

__global__ void k_mb_starts(int E, const unsigned long long *hi, const unsigned long long *lo, int *start) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= E) return;
    start[s] = (s == 0 || hi[s] != hi[s - 1] || lo[s] != lo[s - 1]) ? 1 : 0;
}

// gid: inclusive scan of start.  First and last entry of each key group,
// and the first-entry flag in entry order.
__global__ void k_mb_groups(int E, const unsigned long long *hi, const unsigned long long *lo,
                            const int *start, const int *gid, const int *idx, int *gfirst, int *glast,
                            int *isfirst) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= E) return;
    const int g = gid[s] - 1, e = idx[s];
    isfirst[e] = start[s];
    if (start[s]) gfirst[g] = e;
    if (s == E - 1 || hi[s + 1] != hi[s] || lo[s + 1] != lo[s]) glast[g] = e;
}

__global__ void k_mb_faces(int ng, int ndim, const int *cells, int npc, const int *faces, int fw,
                           const int *gfirst, const int *glast, const int *fid, const int *ent_cell,
                           int *owner, int *neigh, int *fnodes, int ow) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int e0 = gfirst[g], e1 = glast[g];
    const int f = fid[e0];
    owner[f] = ent_cell[e0];
    neigh[f] = e1 != e0 ? ent_cell[e1] : -1;
    int nd[4];
    entry_nodes(e0, ndim, cells, npc, faces, fw, nd);
    for (int i = 0; i < ow; ++i) fnodes[(size_t)f * ow + i] = nd[i];
}

__global__ void k_mb_face_of_entry(int E, const int *idx, const int *gid, const int *gfirst, const int *fid,
                                   int *foe) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < E) foe[idx[s]] = fid[gfirst[gid[s] - 1]];
}

// cell centre: the mean of the cell's nodes (sequential sum, then / count)
__global__ void k_mb_cell_centre(int n, const double *xyz, const int *cells, int cw, double *cc) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double s[3] = {0.0, 0.0, 0.0};
    int cnt = 0;
    for (int k = 0; k < cw; ++k) {
        const int v = cells[(size_t)c * cw + k];
        if (v < 0) continue;
        for (int d = 0; d < 3; ++d) s[d] = cnt == 0 ? xyz[(size_t)v * 3 + d] : s[d] + xyz[(size_t)v * 3 + d];
        ++cnt;
    }
    for (int d = 0; d < 3; ++d) cc[(size_t)c * 3 + d] = s[d] / (double)cnt;
}

// face centre, normal and area (3D: Newell over the face's nodes; 2D: the
// edge), then the normal turned to point out of the owner
__global__ void k_mb_face_geom(int nf, int ndim, const double *xyz, const int *fnodes, int ow,
                               const int *owner, const double *cc, double *area, double *nrm, double *fc) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    double c[3], n[3], a;
    if (ndim == 2) {
        const double *p0 = xyz + (size_t)fnodes[(size_t)f * ow] * 3;
        const double *p1 = xyz + (size_t)fnodes[(size_t)f * ow + 1] * 3;
        for (int d = 0; d < 3; ++d) c[d] = 0.5 * (p0[d] + p1[d]);
        const double dx = p1[0] - p0[0], dy = p1[1] - p0[1];
        a = sqrt(dx * dx + dy * dy);
        n[0] = dy;
        n[1] = -dx;
        n[2] = 0.0;
        if (a > 1e-15) {
            n[0] = n[0] / a;
            n[1] = n[1] / a;
        }
    } else {
        const double *p[4];
        int k = 0;
        for (int i = 0; i < ow; ++i) {
            const int v = fnodes[(size_t)f * ow + i];
            if (v >= 0) p[k++] = xyz + (size_t)v * 3;
        }
        for (int d = 0; d < 3; ++d) {
            double s = p[0][d];
            for (int i = 1; i < k; ++i) s = s + p[i][d];
            c[d] = s / (double)k;
        }
        // component d: sum_i (p_i[a] - p_j[a]) (p_i[b] + p_j[b]), j = i + 1 mod k
        const int A[3] = {1, 2, 0}, B[3] = {2, 0, 1};
        for (int d = 0; d < 3; ++d) {
            double s = 0.0;
            for (int i = 0; i < k; ++i) {
                const double *q = p[i], *r = p[i + 1 == k ? 0 : i + 1];
                const double t = (q[A[d]] - r[A[d]]) * (q[B[d]] + r[B[d]]);
                s = i == 0 ? t : s + t;
            }
            n[d] = s;
        }
        a = sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
        if (a > 1e-15)
            for (int d = 0; d < 3; ++d) n[d] = n[d] / a;
        a = a * 0.5;
    }
    const double *o = cc + (size_t)owner[f] * 3;
    const double dot = ((n[0] * (c[0] - o[0])) + n[1] * (c[1] - o[1])) + n[2] * (c[2] - o[2]);
    if (dot < 0)
        for (int d = 0; d < 3; ++d) n[d] = -n[d];
    area[f] = a;
    for (int d = 0; d < 3; ++d) {
        nrm[(size_t)f * 3 + d] = n[d];
        fc[(size_t)f * 3 + d] = c[d];
    }
}

// 3D cell volume: (1/3) |sum over the cell's faces, in its face-list order,
// of s (x_f . n_f) A_f| with s = -1 where n_f points into the cell; the sum
// is a_0 + (a_1 + a_2 + ...), the per-segment order of numpy's reduceat
__global__ void k_mb_volume3(int n, const int *ptr, const int *foe, const double *cc, const double *fc,
                             const double *nrm, const double *area, double *vol) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double a0 = 0.0, rest = 0.0;
    for (int e = ptr[c]; e < ptr[c + 1]; ++e) {
        const int f = foe[e];
        const double *x = fc + (size_t)f * 3, *m = nrm + (size_t)f * 3, *o = cc + (size_t)c * 3;
        const double tdot = ((m[0] * (x[0] - o[0])) + m[1] * (x[1] - o[1])) + m[2] * (x[2] - o[2]);
        const double sgn = tdot < 0 ? -1.0 : 1.0;
        const double xn = ((x[0] * m[0]) + x[1] * m[1]) + x[2] * m[2];
        const double t = (sgn * xn) * area[f];
        if (e == ptr[c]) a0 = t;
        else rest = rest + t;
    }
    double v = ptr[c + 1] > ptr[c] ? a0 + rest : 0.0;
    v = fabs(v) / 3.0;
    vol[c] = v < 1e-30 ? 1e-30 : v;
}

// 2D cell area: |shoelace| / 2, summed in node order from 0
__global__ void k_mb_volume2(int n, const double *xyz, const int *cells, int npc, double *vol) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double s = 0.0;
    for (int k = 0; k < npc; ++k) {
        const double *p = xyz + (size_t)cells[(size_t)c * npc + k] * 3;
        const double *q = xyz + (size_t)cells[(size_t)c * npc + (k + 1 == npc ? 0 : k + 1)] * 3;
        s = s + (p[0] * q[1] - q[0] * p[1]);
    }
    vol[c] = fabs(s) / 2.0;
}

__global__ void k_mb_bkeys(int nb, const int *bnodes, int bw, unsigned long long *hi, unsigned long long *lo,
                           int *idx) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int nd[4] = {-1, -1, -1, -1};
    for (int i = 0; i < bw && i < 4; ++i) nd[i] = bnodes[(size_t)b * bw + i];
    const Key k = make_key(nd, 4);
    hi[b] = k.hi;
    lo[b] = k.lo;
    idx[b] = b;
}

// patch of each boundary face: the last boundary-list entry with its key
// (upper bound in the stably sorted list, then one back); none: `dflt`
__global__ void k_mb_tag(int nf, const int *neigh, const int *fnodes, int ow, int nb,
                         const unsigned long long *bhi, const unsigned long long *blo, const int *bidx,
                         const int *bname, int dflt, int *patch) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    if (neigh[f] >= 0) {
        patch[f] = -1;
        return;
    }
    int nd[4] = {-1, -1, -1, -1};
    for (int i = 0; i < ow; ++i) nd[i] = fnodes[(size_t)f * ow + i];
    const Key k = make_key(nd, 4);
    int lo_i = 0, hi_i = nb;            // first position with key > k
    while (lo_i < hi_i) {
        const int m = (lo_i + hi_i) >> 1;
        const bool le = bhi[m] < k.hi || (bhi[m] == k.hi && blo[m] <= k.lo);
        if (le) lo_i = m + 1;
        else hi_i = m;
    }
    const int p = lo_i - 1;
    patch[f] = (p >= 0 && bhi[p] == k.hi && blo[p] == k.lo) ? bname[bidx[p]] : dflt;
}

// device buffers of one build, freed together
struct Bufs {
    std::vector<void *> ptrs;
    ~Bufs() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    int get(T **p, size_t count) {
        void *q = nullptr;
        KCFD_HIP(hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1)));
        ptrs.push_back(q);
        *p = (T *)q;
        return KCFD_OK;
    }
};

// stable sort of (hi, lo) keys with their int payload, LSD: by lo, then by hi
int sort_keys(Bufs &bf, int n, unsigned long long *hi, unsigned long long *lo, int *idx, hipStream_t s,
              unsigned long long **hi_out, unsigned long long **lo_out, int **idx_out) {
    unsigned long long *ktmp, *hsrt, *lsrt, *hg;
    int *i1, *i2;
    KCFD_CHECK(bf.get(&ktmp, n));
    KCFD_CHECK(bf.get(&hsrt, n));
    KCFD_CHECK(bf.get(&lsrt, n));
    KCFD_CHECK(bf.get(&hg, n));
    KCFD_CHECK(bf.get(&i1, n));
    KCFD_CHECK(bf.get(&i2, n));
    size_t tb1 = 0, tb2 = 0;
    KCFD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, lo, ktmp, idx, i1, n, 0, 64, s));
    KCFD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, hg, hsrt, i1, i2, n, 0, 64, s));
    void *tmp;
    KCFD_CHECK(bf.get((char **)&tmp, std::max(tb1, tb2)));
    size_t tb = std::max(tb1, tb2);
    KCFD_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, lo, ktmp, idx, i1, n, 0, 64, s));
    if (n) hipLaunchKernelGGL(k_mb_gather, dim3(mb_grid(n)), dim3(kMbBlock), 0, s, n, i1, hi, hg);
    tb = std::max(tb1, tb2);
    KCFD_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, hg, hsrt, i1, i2, n, 0, 64, s));
    if (n) hipLaunchKernelGGL(k_mb_gather, dim3(mb_grid(n)), dim3(kMbBlock), 0, s, n, i2, lo, lsrt);
    KCFD_HIP(hipGetLastError());
    *hi_out = hsrt;
    *lo_out = lsrt;
    *idx_out = i2;
    return KCFD_OK;
}

}  // namespace

struct kcfd_mesh_faces {
    int ndim = 0, n_cells = 0, n_faces = 0, n_internal = 0, face_width = 0, n_nodes = 0;
    std::vector<int> owner, neighbour, face_nodes, face_patch;
    std::vector<double> area, normal, centre, cell_centre, cell_volume, nodes;
    double ms = 0.0;
};

namespace {

// The build from device-resident inputs: xyz (n_nodes x 3), cells
// (n x cw), 3D face lists (E x fw, per-cell offsets ptr), boundary rows
// (nb x fw) and their patch indices.  Results go to r (host).  The stream's
// event ev0 was recorded before the inputs were written.
int build_dev(kcfd_ctx *ctx, Bufs &bf, int ndim, int n_nodes, double *xyz, int n, int cell_width, int *cells,
              long long E, int fw, int *faces, int *ptr, int nb, int *bn, int *bname, int n_patches,
              kcfd_mesh_faces *r) {
    hipStream_t s = ctx->stream;
    r->ndim = ndim;
    r->n_cells = n;
    r->face_width = fw;
    r->n_nodes = n_nodes;
    const int En = (int)E;
    int *ent_cell;
    KCFD_CHECK(bf.get(&ent_cell, E));
    if (n) hipLaunchKernelGGL(k_mb_entry_cell, dim3(mb_grid(n)), dim3(kMbBlock), 0, s, n, ptr, ent_cell);
    unsigned long long *hi, *lo, *hs, *ls;
    int *idx, *is;
    KCFD_CHECK(bf.get(&hi, E));
    KCFD_CHECK(bf.get(&lo, E));
    KCFD_CHECK(bf.get(&idx, E));
    int ng = 0;
    int *gfirst = nullptr, *glast = nullptr, *fid = nullptr, *foe = nullptr, *gid = nullptr;
    if (En) {
        hipLaunchKernelGGL(k_mb_keys, dim3(mb_grid(En)), dim3(kMbBlock), 0, s, En, ndim, cells, cell_width, faces,
                           fw, hi, lo, idx);
        KCFD_CHECK(sort_keys(bf, En, hi, lo, idx, s, &hs, &ls, &is));
        int *start, *isfirst;
        KCFD_CHECK(bf.get(&start, E));
        KCFD_CHECK(bf.get(&gid, E));
        KCFD_CHECK(bf.get(&gfirst, E));
        KCFD_CHECK(bf.get(&glast, E));
        KCFD_CHECK(bf.get(&isfirst, E));
        KCFD_CHECK(bf.get(&fid, E));
        KCFD_CHECK(bf.get(&foe, E));
        hipLaunchKernelGGL(k_mb_starts, dim3(mb_grid(En)), dim3(kMbBlock), 0, s, En, hs, ls, start);
        size_t t1 = 0, t2 = 0;
        KCFD_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, t1, start, gid, En, s));
        KCFD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, isfirst, fid, En, s));
        void *tmp;
        KCFD_CHECK(bf.get((char **)&tmp, std::max(t1, t2)));
        size_t tb = std::max(t1, t2);
        KCFD_HIP(hipcub::DeviceScan::InclusiveSum(tmp, tb, start, gid, En, s));
        hipLaunchKernelGGL(k_mb_groups, dim3(mb_grid(En)), dim3(kMbBlock), 0, s, En, hs, ls, start, gid, is, gfirst,
                           glast, isfirst);
        tb = std::max(t1, t2);
        KCFD_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, isfirst, fid, En, s));
        KCFD_HIP(hipMemcpyAsync(&ng, gid + (En - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        KCFD_HIP(hipStreamSynchronize(s));
    }
    const int nf = ng;
    r->n_faces = nf;
    int *owner, *neigh, *fnodes, *patch;
    double *area, *nrm, *fc, *cc, *vol;
    KCFD_CHECK(bf.get(&owner, nf));
    KCFD_CHECK(bf.get(&neigh, nf));
    KCFD_CHECK(bf.get(&fnodes, (size_t)nf * fw));
    KCFD_CHECK(bf.get(&patch, nf));
    KCFD_CHECK(bf.get(&area, nf));
    KCFD_CHECK(bf.get(&nrm, (size_t)nf * 3));
    KCFD_CHECK(bf.get(&fc, (size_t)nf * 3));
    KCFD_CHECK(bf.get(&cc, (size_t)n * 3));
    KCFD_CHECK(bf.get(&vol, n));
    if (nf) {
        hipLaunchKernelGGL(k_mb_faces, dim3(mb_grid(nf)), dim3(kMbBlock), 0, s, nf, ndim, cells, cell_width, faces,
                           fw, gfirst, glast, fid, ent_cell, owner, neigh, fnodes, fw);
        hipLaunchKernelGGL(k_mb_face_of_entry, dim3(mb_grid(En)), dim3(kMbBlock), 0, s, En, is, gid, gfirst, fid,
                           foe);
    }
    if (n) hipLaunchKernelGGL(k_mb_cell_centre, dim3(mb_grid(n)), dim3(kMbBlock), 0, s, n, xyz, cells, cell_width,
                              cc);
    if (nf) hipLaunchKernelGGL(k_mb_face_geom, dim3(mb_grid(nf)), dim3(kMbBlock), 0, s, nf, ndim, xyz, fnodes, fw,
                               owner, cc, area, nrm, fc);
    if (n) {
        if (ndim == 3)
            hipLaunchKernelGGL(k_mb_volume3, dim3(mb_grid(n)), dim3(kMbBlock), 0, s, n, ptr, foe, cc, fc, nrm, area,
                               vol);
        else
            hipLaunchKernelGGL(k_mb_volume2, dim3(mb_grid(n)), dim3(kMbBlock), 0, s, n, xyz, cells, cell_width, vol);
    }
    // boundary patches
    if (nf) {
        unsigned long long *bhi, *blo, *bhs, *bls;
        int *bidx, *bis;
        KCFD_CHECK(bf.get(&bhi, nb));
        KCFD_CHECK(bf.get(&blo, nb));
        KCFD_CHECK(bf.get(&bidx, nb));
        bhs = bhi;
        bls = blo;
        bis = bidx;
        if (nb) {
            hipLaunchKernelGGL(k_mb_bkeys, dim3(mb_grid(nb)), dim3(kMbBlock), 0, s, nb, bn, fw, bhi, blo, bidx);
            KCFD_CHECK(sort_keys(bf, nb, bhi, blo, bidx, s, &bhs, &bls, &bis));
        }
        hipLaunchKernelGGL(k_mb_tag, dim3(mb_grid(nf)), dim3(kMbBlock), 0, s, nf, neigh, fnodes, fw, nb, bhs, bls,
                           bis, bname, n_patches, patch);
    }
    KCFD_HIP(hipGetLastError());
    KCFD_HIP(hipEventRecord(ctx->ev1, s));
    r->owner.resize(nf);
    r->neighbour.resize(nf);
    r->face_nodes.resize((size_t)nf * fw);
    r->face_patch.resize(nf);
    r->area.resize(nf);
    r->normal.resize((size_t)nf * 3);
    r->centre.resize((size_t)nf * 3);
    r->cell_centre.resize((size_t)n * 3);
    r->cell_volume.resize(n);
    r->nodes.resize((size_t)n_nodes * 3);
    KCFD_HIP(hipMemcpyAsync(r->owner.data(), owner, sizeof(int) * nf, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipMemcpyAsync(r->neighbour.data(), neigh, sizeof(int) * nf, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipMemcpyAsync(r->face_nodes.data(), fnodes, sizeof(int) * (size_t)nf * fw, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipMemcpyAsync(r->face_patch.data(), patch, sizeof(int) * nf, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipMemcpyAsync(r->area.data(), area, sizeof(double) * nf, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipMemcpyAsync(r->normal.data(), nrm, sizeof(double) * (size_t)nf * 3, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipMemcpyAsync(r->centre.data(), fc, sizeof(double) * (size_t)nf * 3, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipMemcpyAsync(r->cell_centre.data(), cc, sizeof(double) * (size_t)n * 3, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipMemcpyAsync(r->cell_volume.data(), vol, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipMemcpyAsync(r->nodes.data(), xyz, sizeof(double) * (size_t)n_nodes * 3, hipMemcpyDeviceToHost, s));
    KCFD_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    KCFD_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    r->ms = ms;
    int ni = 0;
    for (int f = 0; f < nf; ++f) ni += r->neighbour[f] >= 0;
    r->n_internal = ni;
    return KCFD_OK;
}

// ---- structured generators: mesh/mesh_generator.py:13-93 (quads, nodes
// row-major, cells counter-clockwise) and mesh/mesh_generator_3d.py:12-159
// (hexes, node k (ny+1)(nx+1) + j (nx+1) + i, cells i fastest, faces z-, z+,
// y-, y+, x-, x+); node coordinates o + i d, as the generators compute them

__global__ void k_gen_nodes(int nx, int ny, int nz, double ox, double oy, double dx, double dy, double dz,
                            long long nn, double *xyz) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    const long long px = nx + 1, py = ny + 1;
    const int i = (int)(v % px), j = (int)((v / px) % py), k = (int)(v / (px * py));
    xyz[v * 3 + 0] = ox + (double)i * dx;
    xyz[v * 3 + 1] = oy + (double)j * dy;
    xyz[v * 3 + 2] = nz > 0 ? (double)k * dz : 0.0;
}

__global__ void k_gen_quads(int nx, int ny, int *cells) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nx * ny) return;
    const int i = c % nx, j = c / nx, n0 = j * (nx + 1) + i;
    int *o = cells + (size_t)c * 4;
    o[0] = n0;
    o[1] = n0 + 1;
    o[2] = n0 + 1 + (nx + 1);
    o[3] = n0 + (nx + 1);
}

__device__ __forceinline__ int hid(int nx, int ny, int i, int j, int k) {
    return k * (ny + 1) * (nx + 1) + j * (nx + 1) + i;
}

__global__ void k_gen_hexes(int nx, int ny, int nz, int *cells, int *faces, int *ptr) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nx * ny * nz) return;
    const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    const int n[8] = {hid(nx, ny, i, j, k),         hid(nx, ny, i + 1, j, k),
                      hid(nx, ny, i + 1, j + 1, k), hid(nx, ny, i, j + 1, k),
                      hid(nx, ny, i, j, k + 1),     hid(nx, ny, i + 1, j, k + 1),
                      hid(nx, ny, i + 1, j + 1, k + 1), hid(nx, ny, i, j + 1, k + 1)};
    for (int q = 0; q < 8; ++q) cells[(size_t)c * 8 + q] = n[q];
    const int fq[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {2, 3, 7, 6}, {0, 4, 7, 3}, {1, 2, 6, 5}};
    for (int f = 0; f < 6; ++f)
        for (int q = 0; q < 4; ++q) faces[((size_t)c * 6 + f) * 4 + q] = n[fq[f][q]];
    ptr[c] = c * 6;
    if (c == nx * ny * nz - 1) ptr[c + 1] = (c + 1) * 6;
}

// boundary rows of side `side` (quad: bottom, top, left, right; hex: x_min,
// x_max, y_min, y_max, z_min, z_max), row q of it at out[q]
__global__ void k_gen_bside(int ndim, int side, int nx, int ny, int nz, int rows, int patch, int *out, int *pout) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= rows) return;
    int *o = out + (size_t)q * (ndim == 2 ? 2 : 4);
    pout[q] = patch;
    if (ndim == 2) {
        const int w = nx + 1;
        switch (side) {
            case 0: o[0] = q; o[1] = q + 1; break;
            case 1: o[0] = ny * w + q; o[1] = ny * w + q + 1; break;
            case 2: o[0] = q * w; o[1] = (q + 1) * w; break;
            default: o[0] = q * w + nx; o[1] = (q + 1) * w + nx; break;
        }
        return;
    }
    if (side < 2) {                 // (k, j), j fastest
        const int jb = q % ny, kb = q / ny, x = side == 0 ? 0 : nx;
        if (side == 0) {
            o[0] = hid(nx, ny, x, jb, kb); o[1] = hid(nx, ny, x, jb, kb + 1);
            o[2] = hid(nx, ny, x, jb + 1, kb + 1); o[3] = hid(nx, ny, x, jb + 1, kb);
        } else {
            o[0] = hid(nx, ny, x, jb, kb); o[1] = hid(nx, ny, x, jb + 1, kb);
            o[2] = hid(nx, ny, x, jb + 1, kb + 1); o[3] = hid(nx, ny, x, jb, kb + 1);
        }
    } else if (side < 4) {          // (k, i), i fastest
        const int ic = q % nx, kc = q / nx;
        if (side == 2) {
            o[0] = hid(nx, ny, ic, 0, kc); o[1] = hid(nx, ny, ic + 1, 0, kc);
            o[2] = hid(nx, ny, ic + 1, 0, kc + 1); o[3] = hid(nx, ny, ic, 0, kc + 1);
        } else {
            o[0] = hid(nx, ny, ic + 1, ny, kc); o[1] = hid(nx, ny, ic, ny, kc);
            o[2] = hid(nx, ny, ic, ny, kc + 1); o[3] = hid(nx, ny, ic + 1, ny, kc + 1);
        }
    } else {                        // (j, i), i fastest
        const int id = q % nx, jd = q / nx;
        if (side == 4) {
            o[0] = hid(nx, ny, id, jd, 0); o[1] = hid(nx, ny, id, jd + 1, 0);
            o[2] = hid(nx, ny, id + 1, jd + 1, 0); o[3] = hid(nx, ny, id + 1, jd, 0);
        } else {
            o[0] = hid(nx, ny, id, jd, nz); o[1] = hid(nx, ny, id + 1, jd, nz);
            o[2] = hid(nx, ny, id + 1, jd + 1, nz); o[3] = hid(nx, ny, id, jd + 1, nz);
        }
    }
}

}  // namespace

extern "C" int kcfd_mesh_build(kcfd_ctx *ctx, int ndim, int n_nodes, const double *nodes, int n_cells,
                               int cell_width, const int *cell_nodes, int n_entries, int face_width,
                               const int *entry_nodes_h, const int *entry_ptr, int n_bfaces,
                               const int *bface_nodes, const int *bface_patch, int n_patches,
                               kcfd_mesh_faces **out) {
    KCFD_ARG(ctx && out, "null argument");
    KCFD_ARG(ndim == 2 || ndim == 3, "ndim must be 2 or 3");
    KCFD_ARG(n_nodes >= 0 && n_cells >= 0 && n_bfaces >= 0 && n_patches >= 0, "negative size");
    KCFD_ARG(cell_width >= 1 && cell_width <= 8, "cell_width must be 1..8");
    KCFD_ARG(n_cells == 0 || (nodes && cell_nodes), "null mesh array");
    KCFD_ARG(n_bfaces == 0 || (bface_nodes && bface_patch), "null boundary array");
    long long E;
    int fw;
    if (ndim == 2) {
        KCFD_ARG(cell_width >= 3, "a 2D cell needs at least 3 nodes");
        E = (long long)n_cells * cell_width;
        fw = 2;
    } else {
        KCFD_ARG(face_width >= 3 && face_width <= 4, "3D faces have 3 or 4 nodes");
        KCFD_ARG(n_cells == 0 || (entry_nodes_h && entry_ptr), "null face lists");
        E = n_entries;
        fw = face_width;
        KCFD_ARG(n_cells == 0 || entry_ptr[0] == 0, "entry_ptr[0] must be 0");
        for (int c = 0; c < n_cells; ++c) {
            const int k = entry_ptr[c + 1] - entry_ptr[c];
            KCFD_ARG(k >= 1 && k <= 8, "a 3D cell has 1..8 faces");
        }
        KCFD_ARG(n_cells == 0 || entry_ptr[n_cells] == n_entries, "entry_ptr[n] must equal n_entries");
    }
    KCFD_ARG(E < (1LL << 31), "too many cell faces");
    // node ids in range (a key packs them into 31 bits)
    {
        const int *ids = ndim == 2 ? cell_nodes : entry_nodes_h;
        const long long m = ndim == 2 ? (long long)n_cells * cell_width : E * fw;
        for (long long q = 0; q < m; ++q)
            KCFD_ARG(ids[q] < n_nodes && (ids[q] >= 0 || ndim == 3), "node id out of range");
        for (long long q = 0; q < (long long)n_cells * cell_width; ++q)
            KCFD_ARG(cell_nodes[q] < n_nodes, "cell node id out of range");
        const int bw = fw;
        for (long long q = 0; q < (long long)n_bfaces * bw; ++q)
            KCFD_ARG(bface_nodes[q] < n_nodes, "boundary node id out of range");
        // every cell has a node; a 3D face has at least 3 nodes, padding last
        for (int c = 0; c < n_cells; ++c) {
            int cnt = 0;
            for (int k = 0; k < cell_width; ++k) cnt += cell_nodes[(size_t)c * cell_width + k] >= 0;
            KCFD_ARG(cnt >= 1, "a cell without nodes");
        }
        if (ndim == 3)
            for (long long e = 0; e < E; ++e) {
                int cnt = 0;
                bool tail = false, ok = true;
                for (int i = 0; i < fw; ++i) {
                    const bool v = entry_nodes_h[e * fw + i] >= 0;
                    if (v && tail) ok = false;
                    if (!v) tail = true;
                    cnt += v;
                }
                KCFD_ARG(ok && cnt >= 3, "a 3D face needs at least 3 nodes, padding (-1) last");
            }
        for (int b = 0; b < n_bfaces; ++b)
            KCFD_ARG(bface_patch[b] >= 0 && bface_patch[b] < n_patches, "boundary patch index out of range");
    }
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    KCFD_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    std::unique_ptr<kcfd_mesh_faces> r(new kcfd_mesh_faces());
    KCFD_HIP(hipEventRecord(ctx->ev0, s));
    Bufs bf;
    const int n = n_cells;
    double *xyz;
    int *cells, *faces = nullptr, *ptr, *bn, *bname;
    KCFD_CHECK(bf.get(&xyz, (size_t)n_nodes * 3));
    KCFD_CHECK(bf.get(&cells, (size_t)n * cell_width));
    KCFD_CHECK(bf.get(&ptr, (size_t)n + 1));
    KCFD_CHECK(bf.get(&bn, (size_t)n_bfaces * fw));
    KCFD_CHECK(bf.get(&bname, n_bfaces));
    KCFD_HIP(hipMemcpyAsync(xyz, nodes, sizeof(double) * (size_t)n_nodes * 3, hipMemcpyHostToDevice, s));
    KCFD_HIP(hipMemcpyAsync(cells, cell_nodes, sizeof(int) * (size_t)n * cell_width, hipMemcpyHostToDevice, s));
    std::vector<int> hp;
    if (ndim == 3) {
        KCFD_CHECK(bf.get(&faces, E * fw));
        KCFD_HIP(hipMemcpyAsync(faces, entry_nodes_h, sizeof(int) * E * fw, hipMemcpyHostToDevice, s));
        KCFD_HIP(hipMemcpyAsync(ptr, entry_ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice, s));
    } else {
        // 2D: cell c owns entries c * cell_width .. (c + 1) * cell_width - 1
        hp.resize((size_t)n + 1);
        for (int c = 0; c <= n; ++c) hp[c] = c * cell_width;
        KCFD_HIP(hipMemcpyAsync(ptr, hp.data(), sizeof(int) * hp.size(), hipMemcpyHostToDevice, s));
    }
    if (n_bfaces) {
        KCFD_HIP(hipMemcpyAsync(bn, bface_nodes, sizeof(int) * (size_t)n_bfaces * fw, hipMemcpyHostToDevice, s));
        KCFD_HIP(hipMemcpyAsync(bname, bface_patch, sizeof(int) * n_bfaces, hipMemcpyHostToDevice, s));
    }
    KCFD_CHECK(build_dev(ctx, bf, ndim, n_nodes, xyz, n, cell_width, cells, E, fw, faces, ptr, n_bfaces, bn, bname,
                         n_patches, r.get()));
    *out = r.release();
    return KCFD_OK;
}

extern "C" int kcfd_mesh_generate(kcfd_ctx *ctx, int ndim, int nx, int ny, int nz, double ox, double oy, double dx,
                                  double dy, double dz, const int *side_patch, int n_patches,
                                  kcfd_mesh_faces **out) {
    KCFD_ARG(ctx && out && side_patch, "null argument");
    KCFD_ARG(ndim == 2 || ndim == 3, "ndim must be 2 or 3");
    KCFD_ARG(nx >= 1 && ny >= 1 && (ndim == 2 || nz >= 1), "cells per direction must be at least 1");
    const int nsides = ndim == 2 ? 4 : 6;
    for (int q = 0; q < nsides; ++q)
        KCFD_ARG(side_patch[q] >= -1 && side_patch[q] < n_patches, "side patch index out of range");
    const long long nzz = ndim == 3 ? nz : 0;
    const long long nn = (long long)(nx + 1) * (ny + 1) * (nzz + 1);
    const long long n = (long long)nx * ny * (ndim == 3 ? nz : 1);
    const int cw = ndim == 2 ? 4 : 8, fw = ndim == 2 ? 2 : 4;
    const long long E = n * (ndim == 2 ? 4 : 6);
    KCFD_ARG(nn < (1LL << 31) - 1 && E < (1LL << 31), "mesh too large");
    long long rows[6] = {0, 0, 0, 0, 0, 0};
    if (ndim == 2) {
        rows[0] = rows[1] = nx;
        rows[2] = rows[3] = ny;
    } else {
        rows[0] = rows[1] = (long long)ny * nz;
        rows[2] = rows[3] = (long long)nx * nz;
        rows[4] = rows[5] = (long long)nx * ny;
    }
    long long nb = 0;
    for (int q = 0; q < nsides; ++q)
        if (side_patch[q] >= 0) nb += rows[q];
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    KCFD_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    std::unique_ptr<kcfd_mesh_faces> r(new kcfd_mesh_faces());
    KCFD_HIP(hipEventRecord(ctx->ev0, s));
    Bufs bf;
    double *xyz;
    int *cells, *faces = nullptr, *ptr, *bn, *bname;
    KCFD_CHECK(bf.get(&xyz, (size_t)nn * 3));
    KCFD_CHECK(bf.get(&cells, (size_t)n * cw));
    KCFD_CHECK(bf.get(&ptr, (size_t)n + 1));
    KCFD_CHECK(bf.get(&bn, (size_t)nb * fw));
    KCFD_CHECK(bf.get(&bname, nb));
    hipLaunchKernelGGL(k_gen_nodes, dim3(mb_grid(nn)), dim3(kMbBlock), 0, s, nx, ny, (int)nzz, ox, oy, dx, dy, dz, nn,
                       xyz);
    if (ndim == 2) {
        hipLaunchKernelGGL(k_gen_quads, dim3(mb_grid(n)), dim3(kMbBlock), 0, s, nx, ny, cells);
        std::vector<int> hp((size_t)n + 1);
        for (long long c = 0; c <= n; ++c) hp[c] = (int)(c * cw);
        KCFD_HIP(hipMemcpyAsync(ptr, hp.data(), sizeof(int) * hp.size(), hipMemcpyHostToDevice, s));
        KCFD_HIP(hipStreamSynchronize(s));   // hp is a local
    } else {
        KCFD_CHECK(bf.get(&faces, E * fw));
        hipLaunchKernelGGL(k_gen_hexes, dim3(mb_grid(n)), dim3(kMbBlock), 0, s, nx, ny, nz, cells, faces, ptr);
    }
    long long at = 0;
    for (int q = 0; q < nsides; ++q) {
        if (side_patch[q] < 0 || rows[q] == 0) continue;
        hipLaunchKernelGGL(k_gen_bside, dim3(mb_grid(rows[q])), dim3(kMbBlock), 0, s, ndim, q, nx, ny, (int)nzz,
                           (int)rows[q], side_patch[q], bn + at * fw, bname + at);
        at += rows[q];
    }
    KCFD_HIP(hipGetLastError());
    KCFD_CHECK(build_dev(ctx, bf, ndim, (int)nn, xyz, (int)n, cw, cells, E, fw, faces, ptr, (int)nb, bn, bname,
                         n_patches, r.get()));
    *out = r.release();
    return KCFD_OK;
}

extern "C" int kcfd_mesh_faces_nodes(const kcfd_mesh_faces *m, int *n_nodes, int *n_cells, double *nodes) {
    KCFD_ARG(m, "null mesh faces");
    if (n_nodes) *n_nodes = m->n_nodes;
    if (n_cells) *n_cells = m->n_cells;
    if (nodes && !m->nodes.empty()) std::memcpy(nodes, m->nodes.data(), sizeof(double) * m->nodes.size());
    return KCFD_OK;
}

extern "C" int kcfd_mesh_faces_info(const kcfd_mesh_faces *m, int *n_faces, int *n_internal, int *face_width,
                                    double *build_ms) {
    KCFD_ARG(m, "null mesh faces");
    if (n_faces) *n_faces = m->n_faces;
    if (n_internal) *n_internal = m->n_internal;
    if (face_width) *face_width = m->face_width;
    if (build_ms) *build_ms = m->ms;
    return KCFD_OK;
}

extern "C" int kcfd_mesh_faces_get(const kcfd_mesh_faces *m, int *owner, int *neighbour, int *face_nodes,
                                   int *face_patch, double *area, double *normal, double *centre,
                                   double *cell_centre, double *cell_volume) {
    KCFD_ARG(m, "null mesh faces");
    auto cp = [](void *dst, const void *src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    cp(owner, m->owner.data(), sizeof(int) * m->owner.size());
    cp(neighbour, m->neighbour.data(), sizeof(int) * m->neighbour.size());
    cp(face_nodes, m->face_nodes.data(), sizeof(int) * m->face_nodes.size());
    cp(face_patch, m->face_patch.data(), sizeof(int) * m->face_patch.size());
    cp(area, m->area.data(), sizeof(double) * m->area.size());
    cp(normal, m->normal.data(), sizeof(double) * m->normal.size());
    cp(centre, m->centre.data(), sizeof(double) * m->centre.size());
    cp(cell_centre, m->cell_centre.data(), sizeof(double) * m->cell_centre.size());
    cp(cell_volume, m->cell_volume.data(), sizeof(double) * m->cell_volume.size());
    return KCFD_OK;
}

extern "C" int kcfd_mesh_faces_free(kcfd_mesh_faces *m) {
    delete m;
    return KCFD_OK;
}
