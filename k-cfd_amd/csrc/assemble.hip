// Face-parallel FVM assembly into a device-resident CSR matrix, Green-Gauss
// gradients, MUSCL/TVD deferred correction and face mass fluxes (gfx950).
//
// Layout: SoA mesh (owner/neighbour/area/normal/centres per face, centre and
// volume per cell), a cell->face CSR in face-id order with the side of the
// cell (+1 owner, -1 neighbour), and a static face->slot map into the
// canonical CSR (slot_on = position of (O,N), slot_no = position of (N,O),
// diag_slot[c] = position of (c,c)).
//
// Each operator is (a) a face-parallel kernel that computes the face
// coefficient once, adds it to its two off-diagonal slots (a conforming mesh
// gives every slot exactly one face, so there are no atomics and no races)
// and stores the per-face diagonal contributions, then (b) a cell-parallel
// kernel that walks the cell's faces in face-id order and adds the
// contributions to the diagonal slot, the FVMSystem diagonal tracker and the
// rhs sequentially, in the order the reference's per-face loops append them
// (core/fvm_operators.py).  The result is deterministic and follows the
// reference's summation order per cell.
#include "common.hpp"

#include <algorithm>

struct kcfd_mesh {
    kcfd_ctx *ctx = nullptr;
    int n = 0, nf = 0, ndim = 2, npatch = 0, nnz = 0, nbf = 0;
    int *owner = nullptr, *nb = nullptr;
    double *area = nullptr, *normal = nullptr, *fc = nullptr, *cc = nullptr, *vol = nullptr;
    double *dpn = nullptr, *dpf = nullptr, *gw = nullptr;
    int *bpatch = nullptr, *bidx = nullptr;
    int *cf_ptr = nullptr, *cf_face = nullptr;
    signed char *cf_side = nullptr;
    int *slot_on = nullptr, *slot_no = nullptr, *diag_slot = nullptr;
    double *fco = nullptr;     // per-face scratch: diagonal contribution of O (and of N)
    double *fco2 = nullptr;
    int *d_kinds = nullptr;    // per-patch BC kinds
    double *d_vals = nullptr;  // per-patch Neumann values
    std::vector<int> h_indptr, h_indices;
};

namespace kcfd {

struct MeshView {
    int n, nf, ndim;
    const int *owner, *nb;
    const double *area, *normal, *fc, *cc, *vol, *dpn, *dpf, *gw;
    const int *bpatch, *bidx, *cf_ptr, *cf_face;
    const signed char *cf_side;
    const int *slot_on, *slot_no, *diag_slot;
};

static MeshView view(const kcfd_mesh *m) {
    return MeshView{m->n,      m->nf,     m->ndim,    m->owner,   m->nb,     m->area,
                    m->normal, m->fc,     m->cc,      m->vol,     m->dpn,    m->dpf,
                    m->gw,     m->bpatch, m->bidx,    m->cf_ptr,  m->cf_face, m->cf_side,
                    m->slot_on, m->slot_no, m->diag_slot};
}

static inline unsigned blocks(int n) { return (unsigned)std::max(1, (n + kBlock - 1) / kBlock); }

// --------------------------------------------------------- diffusion

// fvm_operators.py:69-95: internal faces with d_PN >= 1e-30
__global__ __launch_bounds__(kBlock) void k_diff_face(MeshView m, const double *__restrict__ gamma,
                                                      double *__restrict__ vals,
                                                      double *__restrict__ fco) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    const int o = m.owner[f], nb = m.nb[f];
    if (nb < 0 || m.dpn[f] < 1e-30 || m.slot_on[f] < 0) { fco[f] = 0.0; return; }
    const double go = gamma[o], gn = gamma[nb];
    double gf = 0.0;
    if (go + gn > 1e-30) gf = 2.0 * go * gn / (go + gn);
    const double c = gf * m.area[f] / m.dpn[f];
    vals[m.slot_on[f]] += -c;
    vals[m.slot_no[f]] += -c;
    fco[f] = c;
}

// internal-face diagonal gather: contribution fco[f] to both sides, in face order
__global__ __launch_bounds__(kBlock) void k_gather_sym(MeshView m, const double *__restrict__ fco,
                                                       double *__restrict__ vals,
                                                       double *__restrict__ diagt) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int ds = m.diag_slot[c];
    double v = vals[ds], t = diagt[c];
    for (int q = m.cf_ptr[c]; q < m.cf_ptr[c + 1]; ++q) {
        const int f = m.cf_face[q];
        if (m.nb[f] < 0 || m.slot_on[f] < 0) continue;
        const double a = fco[f];
        v += a;
        t += a;
    }
    vals[ds] = v;
    diagt[c] = t;
}

// ------------------------------------------------------------ upwind

// fvm_operators.py:113-124 off-diagonals: (O,N) += min(F,0), (N,O) += min(-F,0)
__global__ __launch_bounds__(kBlock) void k_upw_face(MeshView m, const double *__restrict__ F,
                                                     double *__restrict__ vals) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    if (m.nb[f] < 0) return;
    const double q = F[f];
    vals[m.slot_on[f]] += fmin(q, 0.0);
    vals[m.slot_no[f]] += fmin(-q, 0.0);
}

// diagonal: owner side max(F,0) (boundary: F if F >= 0), neighbour side max(-F,0)
__global__ __launch_bounds__(kBlock) void k_upw_cell(MeshView m, const double *__restrict__ F,
                                                     double *__restrict__ vals,
                                                     double *__restrict__ diagt) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int ds = m.diag_slot[c];
    double v = vals[ds], t = diagt[c];
    for (int q = m.cf_ptr[c]; q < m.cf_ptr[c + 1]; ++q) {
        const int f = m.cf_face[q];
        const double fl = F[f];
        double a;
        if (m.nb[f] >= 0) {
            a = m.cf_side[q] > 0 ? fmax(fl, 0.0) : fmax(-fl, 0.0);
        } else {
            if (!(fl >= 0.0)) continue;
            a = fl;
        }
        v += a;
        t += a;
    }
    vals[ds] = v;
    diagt[c] = t;
}

// -------------------------------------------------------- boundaries
// fvm_operators.py:189-225.  A cell's boundary faces are applied in patch
// order, then patch-local order (= ascending bidx), as the reference's
// patch-by-patch loop does.
__global__ __launch_bounds__(kBlock) void k_bc_cell(MeshView m, const double *__restrict__ gamma,
                                                    const double *__restrict__ F,
                                                    const double *__restrict__ phib,
                                                    const int *__restrict__ kinds,
                                                    const double *__restrict__ pvals,
                                                    double *__restrict__ vals,
                                                    double *__restrict__ diagt,
                                                    double *__restrict__ rhs) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int ds = m.diag_slot[c];
    double v = vals[ds], t = diagt[c], r = rhs[c];
    int last = -1;
    for (;;) {
        int best = -1, bf = -1;
        for (int q = m.cf_ptr[c]; q < m.cf_ptr[c + 1]; ++q) {
            const int f = m.cf_face[q];
            const int bi = m.bidx[f];
            if (m.nb[f] >= 0 || bi < 0 || bi <= last) continue;
            if (best < 0 || bi < best) { best = bi; bf = f; }
        }
        if (best < 0) break;
        last = best;
        const int f = bf;
        if (m.dpf[f] < 1e-30) continue;
        const int kind = kinds[m.bpatch[f]];
        const double fl = F[f];
        if (kind == KCFD_BC_DIRICHLET) {
            const double pb = phib[best];
            const double cf = gamma[c] * m.area[f] / m.dpf[f];
            v += cf;
            t += cf;
            r += cf * pb;
            if (fl < 0.0) r += -fl * pb;
        } else if (kind == KCFD_BC_NEUMANN) {
            r += pvals[m.bpatch[f]] * m.area[f];
        } else if (kind == KCFD_BC_ZERO_GRADIENT) {
            if (fl < 0.0) {
                v += -fl;
                t += -fl;
            }
        }
    }
    vals[ds] = v;
    diagt[c] = t;
    rhs[c] = r;
}

// ------------------------------------------------------- cell terms
// mode 0 temporal (fvm_operators.py:144-148): k = rho*V/dt; diag += k; rhs += k*phi_old
// mode 1 source   (:159-161): rhs += S*V
// mode 2 linearised (:172-175): diag += -Sp*V; rhs += Su*V
// mode 3 drag     (models/two_fluid.py:592-596): diag += K*V; rhs += (K*V)*u_other
//        (passed as kd = K, ks = u_other)
// mode 4 under-relax (:234-240): ap = tracker; diag += ap*(1-a)/a; rhs += ap*(1-a)/a*phi
__global__ __launch_bounds__(kBlock) void k_cell_terms(int n, int mode, const int *__restrict__ dslot,
                                                       const double *__restrict__ vol,
                                                       const double *__restrict__ a1,
                                                       const double *__restrict__ a2, double s1,
                                                       double s2, double *__restrict__ vals,
                                                       double *__restrict__ diagt,
                                                       double *__restrict__ rhs) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int ds = dslot[c];
    if (mode == 0) {
        const double k = s1 * vol[c] / s2;
        vals[ds] += k;
        diagt[c] += k;
        rhs[c] += k * a1[c];
    } else if (mode == 1) {
        rhs[c] += a1[c] * vol[c];
    } else if (mode == 2) {
        const double k = -a1[c] * vol[c];
        vals[ds] += k;
        diagt[c] += k;
        rhs[c] += a2[c] * vol[c];
    } else if (mode == 3) {
        const double k = a1[c] * vol[c];
        vals[ds] += k;
        diagt[c] += k;
        rhs[c] += k * a2[c];
    } else {
        const double ap = diagt[c];
        const double k = ap * (1.0 - s1) / s1;
        vals[ds] += k;
        diagt[c] += k;
        rhs[c] += ap * (1.0 - s1) / s1 * a1[c];
    }
}

// ---------------------------------------------- interphase drag closure
// schiller_naumann_drag (models/closure.py:13-34): Re floored at 1e-10 (NaN
// kept, as np.maximum does), C_D = 24/Re (1 + 0.15 Re^0.687) below 1000,
// else 0.44; operation order as the numpy expression.
__device__ __forceinline__ double schiller_naumann(double re) {
#pragma clang fp contract(off)
    re = (re >= 1e-10 || re != re) ? re : 1e-10;
    return re < 1000.0 ? 24.0 / re * (1.0 + 0.15 * pow(re, 0.687)) : 0.44;
}

__global__ __launch_bounds__(kBlock) void k_schiller_naumann(const double *__restrict__ re,
                                                             double *__restrict__ cd, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cd[i] = schiller_naumann(re[i]);
}

// drag_coefficient_implicit (models/closure.py:72-92): |u_rel| (squares
// summed over the ndim components in order, floored at 1e-15),
// Re = rho_l |u_rel| d_b / mu_l, K = 0.75 C_D alpha_g rho_l |u_rel| / d_b.
// u_g, u_l: n x ndim row-major.
__global__ __launch_bounds__(kBlock) void k_drag(int n, int ndim, const double *__restrict__ alpha_g, double rho_l,
                                                 const double *__restrict__ ug, const double *__restrict__ ul,
                                                 double d_b, double mu_l, double *__restrict__ K) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double s = 0.0;
    for (int k = 0; k < ndim; ++k) {
        const double u = ug[(size_t)c * ndim + k] - ul[(size_t)c * ndim + k];
        const double u2 = u * u;
        s = k == 0 ? u2 : s + u2;
    }
    double mag = sqrt(s);
    mag = (mag >= 1e-15 || mag != mag) ? mag : 1e-15;
    const double re = rho_l * mag * d_b / mu_l;
    const double cd = schiller_naumann(re);
    K[c] = 0.75 * cd * alpha_g[c] * rho_l * mag / d_b;
}

// ---------------------------------------------- pressure correction
// models/single_phase.py:288-325 coefficients
__global__ __launch_bounds__(kBlock) void k_press_face(MeshView m, const double *__restrict__ dP,
                                                       double rho, const int *__restrict__ pdir,
                                                       double *__restrict__ vals,
                                                       double *__restrict__ fco) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    const int o = m.owner[f], nb = m.nb[f];
    if (nb >= 0) {
        const double d = m.dpn[f];
        if (d < 1e-30 || m.slot_on[f] < 0) { fco[f] = 0.0; return; }
        const double dx = m.cc[3 * nb] - m.cc[3 * o];
        const double dy = m.cc[3 * nb + 1] - m.cc[3 * o + 1];
        const double dz = m.cc[3 * nb + 2] - m.cc[3 * o + 2];
        double dorth = fabs(dx * m.normal[3 * f] + dy * m.normal[3 * f + 1] + dz * m.normal[3 * f + 2]);
        dorth = fmax(dorth, 0.1 * d);
        const double df = rho * 0.5 * (dP[o] + dP[nb]) * m.area[f] / dorth;
        vals[m.slot_on[f]] += -df;
        vals[m.slot_no[f]] += -df;
        fco[f] = df;
    } else {
        double df = 0.0;
        const int p = m.bpatch[f];
        if (p >= 0 && pdir[p]) {
            const double d = m.dpf[f];
            if (d > 1e-30) df = rho * dP[o] * m.area[f] / d;
        }
        fco[f] = df;
    }
}

// diagonal = running sum of the face coefficients in face order (internal
// faces both sides, Dirichlet boundary faces the owner)
__global__ __launch_bounds__(kBlock) void k_press_cell(MeshView m, const double *__restrict__ fco,
                                                       double *__restrict__ vals,
                                                       double *__restrict__ diagt) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int ds = m.diag_slot[c];
    double a = 0.0;
    for (int q = m.cf_ptr[c]; q < m.cf_ptr[c + 1]; ++q) a += fco[m.cf_face[q]];
    vals[ds] += a;
    diagt[c] += a;
}

// ------------------------------------------ SIMPLE pressure-correction step
// models/single_phase.py:274-365 around the solve.

// dP = V / max((aP_0 + aP_1 [+ aP_2]) / ndim, 1e-30) (the momentum diagonals,
// summed in component order, :279-283)
__global__ __launch_bounds__(kBlock) void k_simple_dp(int n, int nd, const double *__restrict__ vol,
                                                      const double *__restrict__ ap0,
                                                      const double *__restrict__ ap1,
                                                      const double *__restrict__ ap2,
                                                      double *__restrict__ dP) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double s = ap0[c] + ap1[c];
    if (nd == 3) s = s + ap2[c];
    dP[c] = vol[c] / fmax(s / (double)nd, 1e-30);
}

// mass imbalance (:305-311): b_O -= F_f, b_N += F_f, boundary b_O -= F_f, in
// face order per cell
__global__ __launch_bounds__(kBlock) void k_simple_rhs(MeshView m, const double *__restrict__ mf,
                                                       double *__restrict__ rhs) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    double b = rhs[c];
    for (int q = m.cf_ptr[c]; q < m.cf_ptr[c + 1]; ++q) {
        const double F = mf[m.cf_face[q]];
        if (m.cf_side[q] > 0) b -= F;
        else b += F;
    }
    rhs[c] = b;
}

// no Dirichlet p' patch (:327-343): row 0's entries become 0, then the
// appended (0, 0, 1.0) entry lands on its diagonal; b_0 = 0
__global__ void k_simple_fix_row0(const int *__restrict__ ip, const int *__restrict__ dpos,
                                  double *__restrict__ vals, double *__restrict__ rhs) {
    const int q = ip[0] + (int)threadIdx.x;
    if (q < ip[1]) vals[q] = q == dpos[0] ? 0.0 + 1.0 : 0.0;
    if (threadIdx.x == 0) rhs[0] = 0.0;
}

// velocity and pressure correction (:345-363): grad p' by Green-Gauss with
// the boundary values 0 on Dirichlet p' patches and the owner's p' elsewhere
// (faces in no patch take the owner value too), then U -= dP grad p' per
// component and p += alpha_p p'.  U is (n, ndim) row-major.
__global__ __launch_bounds__(kBlock) void k_simple_correct(MeshView m, const double *__restrict__ pp,
                                                           const double *__restrict__ dP,
                                                           const int *__restrict__ pdir, double alpha_p,
                                                           double *__restrict__ U, double *__restrict__ p) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int nd = m.ndim;
    double g[3] = {0.0, 0.0, 0.0};
    for (int q = m.cf_ptr[c]; q < m.cf_ptr[c + 1]; ++q) {
        const int f = m.cf_face[q];
        const int o = m.owner[f], nb = m.nb[f];
        double pf;
        if (nb >= 0) {
            const double w = m.gw[f];
            pf = w * pp[o] + (1.0 - w) * pp[nb];
        } else {
            const int pt = m.bpatch[f];
            pf = (pt >= 0 && pdir[pt]) ? 0.0 : pp[o];
        }
        const double A = m.area[f];
        for (int k = 0; k < 3; ++k) {
            const double fl = pf * m.normal[3 * f + k] * A;
            if (m.cf_side[q] > 0) g[k] += fl;
            else g[k] -= fl;
        }
    }
    const double V = m.vol[c];
    if (V > 1e-30)
        for (int k = 0; k < 3; ++k) g[k] /= V;
    for (int k = 0; k < nd; ++k) U[(size_t)c * nd + k] -= dP[c] * g[k];
    p[c] += alpha_p * pp[c];
}

// ------------------------------ two-fluid pressure-correction step
// models/two_fluid.py:641-725.

// Per face (:655-688): coefficient (rho_l al_f V_f / max(aPl_f, 1e-30) +
// rho_g ag_f V_f / max(aPg_f, 1e-30)) A / d_PN with face averages
// 0.5 (x_O + x_N), written to both off-diagonal slots; the face's total mass
// flux mf_l al_O + mf_g ag_O for the imbalance.  A face with d_PN < 1e-30
// contributes nothing (the reference skips it whole); boundary faces add
// only their mass flux.
__global__ __launch_bounds__(kBlock) void k_tf_press_face(MeshView m, const double *__restrict__ al,
                                                          const double *__restrict__ ag,
                                                          const double *__restrict__ apl,
                                                          const double *__restrict__ apg, double rho_l,
                                                          double rho_g, const double *__restrict__ mfl,
                                                          const double *__restrict__ mfg,
                                                          double *__restrict__ vals,
                                                          double *__restrict__ fco,
                                                          double *__restrict__ tmf) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    const int o = m.owner[f], nb = m.nb[f];
    const double t = mfl[f] * al[o] + mfg[f] * ag[o];
    if (nb >= 0) {
        const double d = m.dpn[f];
        if (d < 1e-30 || m.slot_on[f] < 0) {
            fco[f] = 0.0;
            tmf[f] = 0.0;
            return;
        }
        const double alf = 0.5 * (al[o] + al[nb]);
        const double agf = 0.5 * (ag[o] + ag[nb]);
        const double aplf = 0.5 * (apl[o] + apl[nb]);
        const double apgf = 0.5 * (apg[o] + apg[nb]);
        const double vf = 0.5 * (m.vol[o] + m.vol[nb]);
        const double cl = rho_l * alf * vf / fmax(aplf, 1e-30);
        const double cg = rho_g * agf * vf / fmax(apgf, 1e-30);
        const double c = (cl + cg) * m.area[f] / d;
        vals[m.slot_on[f]] += -c;
        vals[m.slot_no[f]] += -c;
        fco[f] = c;
    } else {
        fco[f] = 0.0;
    }
    tmf[f] = t;
}

// the reference cell (:690): diag_0 += 1e10 after every face
__global__ void k_tf_ref_cell(const int *__restrict__ dpos, double *__restrict__ vals,
                              double *__restrict__ diagt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        vals[dpos[0]] += 1e10;
        diagt[0] += 1e10;
    }
}

// Velocity and pressure correction (:699-723): g = sum over the cell's
// internal faces, in face order, of +-(p'_N - p'_O) n A (owner +, neighbour
// -), divided by V when V > 1e-30; then for each phase
// U_k += alpha_p (-alpha_k V / max(aP_k, 1e-20)) g, and p += alpha_p p'.
// U_l and U_g are (n, ndim) row-major.
__global__ __launch_bounds__(kBlock) void k_tf_correct(MeshView m, const double *__restrict__ pp,
                                                       const double *__restrict__ al,
                                                       const double *__restrict__ ag,
                                                       const double *__restrict__ apl,
                                                       const double *__restrict__ apg, double alpha_p,
                                                       double *__restrict__ Ul, double *__restrict__ Ug,
                                                       double *__restrict__ p) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int nd = m.ndim;
    double g[3] = {0.0, 0.0, 0.0};
    for (int q = m.cf_ptr[c]; q < m.cf_ptr[c + 1]; ++q) {
        const int f = m.cf_face[q];
        const int nb = m.nb[f];
        if (nb < 0) continue;
        const double dp = pp[nb] - pp[m.owner[f]];
        const double A = m.area[f];
        for (int k = 0; k < nd; ++k) {
            const double fl = dp * m.normal[3 * f + k] * A;
            if (m.cf_side[q] > 0) g[k] += fl;
            else g[k] -= fl;
        }
    }
    const double V = m.vol[c];
    if (V > 1e-30)
        for (int k = 0; k < nd; ++k) g[k] /= V;
    const double kl = -al[c] * V / fmax(apl[c], 1e-20);
    const double kg = -ag[c] * V / fmax(apg[c], 1e-20);
    for (int k = 0; k < nd; ++k) {
        Ul[(size_t)c * nd + k] += alpha_p * (kl * g[k]);
        Ug[(size_t)c * nd + k] += alpha_p * (kg * g[k]);
    }
    p[c] += alpha_p * pp[c];
}

// ------------------------------------------------------- Green-Gauss
// core/gradient.py:12-48
__global__ __launch_bounds__(kBlock) void k_gg(MeshView m, const double *__restrict__ phi,
                                               const double *__restrict__ phib,
                                               double *__restrict__ grad) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int nd = m.ndim;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int q = m.cf_ptr[c]; q < m.cf_ptr[c + 1]; ++q) {
        const int f = m.cf_face[q];
        const int o = m.owner[f], nb = m.nb[f];
        double pf;
        if (nb >= 0) {
            const double w = m.gw[f];
            pf = w * phi[o] + (1.0 - w) * phi[nb];
        } else {
            const int bi = m.bidx[f];
            pf = bi >= 0 ? phib[bi] : phi[o];
        }
        const double A = m.area[f];
        const double f0 = pf * m.normal[3 * f] * A;
        const double f1 = pf * m.normal[3 * f + 1] * A;
        const double f2 = pf * m.normal[3 * f + 2] * A;
        if (m.cf_side[q] > 0) {
            g0 += f0; g1 += f1; g2 += f2;
        } else {
            g0 -= f0; g1 -= f1; g2 -= f2;
        }
    }
    const double V = m.vol[c];
    if (V > 1e-30) {
        g0 /= V; g1 /= V; g2 /= V;
    }
    grad[(size_t)c * nd] = g0;
    grad[(size_t)c * nd + 1] = g1;
    if (nd == 3) grad[(size_t)c * nd + 2] = g2;
}

// ------------------------------------------------------------ MUSCL
// core/interpolation.py:227-303 (r uses x_f - x_U, as the reference)
__device__ __forceinline__ double limiter(int kind, double r) {
#pragma clang fp contract(off)
    switch (kind) {
        case KCFD_LIM_MINMOD: return fmax(0.0, fmin(1.0, r));
        case KCFD_LIM_SUPERBEE: return fmax(0.0, fmax(fmin(2.0 * r, 1.0), fmin(r, 2.0)));
        case KCFD_LIM_VAN_ALBADA: return r > 0 ? (r * r + r) / (r * r + 1.0) : 0.0;
        default: return r > 0 ? 2.0 * r / (1.0 + r) : 0.0;
    }
}

// Face values of core/interpolation.py:55-161.  mode 0: upwind by the sign
// of F (F >= 0: the owner's value); 1: central g phi_O + (1 - g) phi_N with
// the owner weight g = d_N / (d_O + d_N) (0.5 when the sum < 1e-30); 2: the
// per-face van Leer TVD value of tvd_face_value, phi_U + 0.5 psi (phi_D -
// phi_U) with psi = (r + |r|) / (1 + |r|), r = 2 (phi_lin - phi_U) /
// (phi_D - phi_U) - 1, phi_lin = phi_U + grad_U . (x_f - x_U), and phi_U
// when |phi_D - phi_U| < 1e-15.  Boundary faces take the patch value (the
// owner's value for a face in no patch).
__global__ __launch_bounds__(kBlock) void k_face_values(MeshView m, int mode, const double *__restrict__ phi,
                                                        const double *__restrict__ phib,
                                                        const double *__restrict__ F,
                                                        const double *__restrict__ grad,
                                                        double *__restrict__ phif) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    const int o = m.owner[f], nb = m.nb[f];
    if (nb < 0) {
        const int bi = m.bidx[f];
        phif[f] = bi >= 0 ? phib[bi] : phi[o];
        return;
    }
    if (mode == 1) {
        const double w = m.gw[f];
        phif[f] = w * phi[o] + (1.0 - w) * phi[nb];
        return;
    }
    const bool pos = F[f] >= 0;
    const int up = pos ? o : nb, dn = pos ? nb : o;
    const double pu = phi[up];
    if (mode == 0) {
        phif[f] = pu;
        return;
    }
    const int nd = m.ndim;
    double dot = grad[(size_t)up * nd] * (m.fc[3 * f] - m.cc[3 * up]);
    dot = dot + grad[(size_t)up * nd + 1] * (m.fc[3 * f + 1] - m.cc[3 * up + 1]);
    if (nd == 3) dot = dot + grad[(size_t)up * nd + 2] * (m.fc[3 * f + 2] - m.cc[3 * up + 2]);
    const double lin = pu + dot;
    const double dup = lin - pu;
    const double dtot = phi[dn] - pu;
    if (fabs(dtot) < 1e-15) {
        phif[f] = pu;
        return;
    }
    const double r = 2.0 * dup / dtot - 1.0;
    const double ar = fabs(r);
    const double psi = (1.0 + ar) > 1e-15 ? (r + ar) / (1.0 + ar) : 0.0;
    phif[f] = pu + 0.5 * psi * dtot;
}

// psi = limiter(r) elementwise (interpolation.py:17-44)
__global__ void k_limiter(int kind, const double *__restrict__ r, double *__restrict__ psi, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) psi[i] = limiter(kind, r[i]);
}

// per internal face: contribution to the owner's correction (the
// neighbour receives its negation)
// core/gradient.py:51-99: least-squares gradient.  Per cell, over its faces
// (face-id order): dx = x_N - x_P, or x_f - x_P on a boundary face; dphi =
// phi_N - phi_P, or the patch value (owner value if none) - phi_P; then
// (A^T A) g = A^T dphi solved by LU with partial pivoting (LAPACK gesv's
// algorithm: first largest |pivot|, row swaps, forward and back
// substitution).  An exactly zero pivot (LAPACK's singular report) or fewer
// faces than dimensions leaves g = 0, as the reference does.
__global__ __launch_bounds__(kBlock) void k_lsq(MeshView m, const double *__restrict__ phi,
                                                const double *__restrict__ phib,
                                                double *__restrict__ grad) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int nd = m.ndim;
    double a[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double r[3] = {0.0, 0.0, 0.0};
    const double pc = phi[c];
    const int q0 = m.cf_ptr[c], q1 = m.cf_ptr[c + 1];
    for (int q = q0; q < q1; ++q) {
        const int f = m.cf_face[q];
        const int o = m.owner[f], nb = m.nb[f];
        double dx[3], dp;
        if (nb >= 0) {
            const int other = m.cf_side[q] > 0 ? nb : o;
            for (int k = 0; k < 3; ++k) dx[k] = m.cc[3 * other + k] - m.cc[3 * c + k];
            dp = phi[other] - pc;
        } else {
            const int bi = m.bidx[f];
            for (int k = 0; k < 3; ++k) dx[k] = m.fc[3 * f + k] - m.cc[3 * c + k];
            dp = (bi >= 0 ? phib[bi] : phi[o]) - pc;
        }
        for (int i = 0; i < nd; ++i) {
            for (int j = 0; j < nd; ++j) a[i][j] += dx[i] * dx[j];
            r[i] += dx[i] * dp;
        }
    }
    double g[3] = {0.0, 0.0, 0.0};
    bool ok = q1 - q0 >= nd;
    // LU with partial pivoting, in place
    int piv[3] = {0, 1, 2};
    for (int k = 0; k < nd && ok; ++k) {
        int p = k;
        for (int i = k + 1; i < nd; ++i)
            if (fabs(a[i][k]) > fabs(a[p][k])) p = i;
        if (a[p][k] == 0.0) { ok = false; break; }
        if (p != k) {
            for (int j = 0; j < nd; ++j) {
                const double t = a[k][j]; a[k][j] = a[p][j]; a[p][j] = t;
            }
            const int t = piv[k]; piv[k] = piv[p]; piv[p] = t;
        }
        for (int i = k + 1; i < nd; ++i) {
            a[i][k] = a[i][k] / a[k][k];
            for (int j = k + 1; j < nd; ++j) a[i][j] -= a[i][k] * a[k][j];
        }
    }
    if (ok) {
        double y[3];
        for (int i = 0; i < nd; ++i) {      // L y = P r
            double s = r[piv[i]];
            for (int j = 0; j < i; ++j) s -= a[i][j] * y[j];
            y[i] = s;
        }
        for (int i = nd - 1; i >= 0; --i) { // U g = y
            double s = y[i];
            for (int j = i + 1; j < nd; ++j) s -= a[i][j] * g[j];
            g[i] = s / a[i][i];
        }
    }
    for (int k = 0; k < nd; ++k) grad[(size_t)c * nd + k] = g[k];
}

__global__ __launch_bounds__(kBlock) void k_muscl_face(MeshView m, const double *__restrict__ phi,
                                                       const double *__restrict__ F,
                                                       const double *__restrict__ grad, int lim,
                                                       double *__restrict__ fco,
                                                       double *__restrict__ fco2) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    const int o = m.owner[f], nb = m.nb[f];
    if (nb < 0) return;
    const double q = F[f];
    const bool pos = q >= 0;
    const int up = pos ? o : nb, dn = pos ? nb : o;
    const double pu = phi[up], pd = phi[dn];
    const int nd = m.ndim;
    double dot = grad[(size_t)up * nd] * (m.fc[3 * f] - m.cc[3 * up]);
    dot = dot + grad[(size_t)up * nd + 1] * (m.fc[3 * f + 1] - m.cc[3 * up + 1]);
    if (nd == 3) dot = dot + grad[(size_t)up * nd + 2] * (m.fc[3 * f + 2] - m.cc[3 * up + 2]);
    const double lin = pu + dot;
    const double dface = lin - pu;
    const double dcd = pd - pu;
    double co = 0.0, cn = 0.0;   // owner, neighbour contributions
    bool act = !(fabs(dcd) < 1e-15);
    if (act) {
        const double r = 2.0 * dface / dcd - 1.0;
        const double psi = limiter(lim, r);
        const double fc = fabs(q) * ((pu + 0.5 * psi * dcd) - pu);
        if (pos) { co = -fc; cn = fc; } else { cn = -fc; co = fc; }
    }
    fco[f] = act ? co : __builtin_nan("");
    fco2[f] = cn;
}

__global__ __launch_bounds__(kBlock) void k_muscl_cell(MeshView m, const double *__restrict__ fco,
                                                       const double *__restrict__ fco2,
                                                       double *__restrict__ corr) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    double s = 0.0;
    for (int q = m.cf_ptr[c]; q < m.cf_ptr[c + 1]; ++q) {
        const int f = m.cf_face[q];
        if (m.nb[f] < 0) continue;
        const double a = fco[f];
        if (a != a) continue;     // face skipped (|phi_D - phi_U| < 1e-15)
        s += m.cf_side[q] > 0 ? a : fco2[f];
    }
    corr[c] = s;
}

// --------------------------------------------------------- mass flux
// core/interpolation.py:163-182
__global__ __launch_bounds__(kBlock) void k_mflux(MeshView m, const double *__restrict__ U,
                                                  const double *__restrict__ Ub, double rho,
                                                  double *__restrict__ F) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    const int nd = m.ndim;
    const int o = m.owner[f], nb = m.nb[f];
    double u[3] = {0.0, 0.0, 0.0};
    if (nb >= 0) {
        const double w = m.gw[f];
        for (int k = 0; k < nd; ++k) u[k] = w * U[(size_t)o * nd + k] + (1.0 - w) * U[(size_t)nb * nd + k];
    } else {
        const int bi = m.bidx[f];
        for (int k = 0; k < nd; ++k) u[k] = bi >= 0 ? Ub[(size_t)bi * nd + k] : U[(size_t)o * nd + k];
    }
    double dot = u[0] * m.normal[3 * f];
    dot = dot + u[1] * m.normal[3 * f + 1];
    if (nd == 3) dot = dot + u[2] * m.normal[3 * f + 2];
    F[f] = rho * dot * m.area[f];
}

__global__ void k_zero(double *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = 0.0;
}

}  // namespace kcfd

using namespace kcfd;

namespace {

template <typename T>
int upload(T **dst, const T *src, size_t count, hipStream_t s) {
    KCFD_HIP(hipMalloc(dst, sizeof(T) * std::max<size_t>(count, 1)));
    if (count && src) KCFD_HIP(hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
    return KCFD_OK;
}

struct Timer {
    kcfd_ctx *ctx;
    explicit Timer(kcfd_ctx *c) : ctx(c) { (void)hipEventRecord(ctx->ev0, ctx->stream); }
    int stop() {
        KCFD_HIP(hipGetLastError());
        KCFD_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        return KCFD_OK;
    }
};

}  // namespace

extern "C" {

int kcfd_mesh_create(kcfd_ctx *ctx, const kcfd_mesh_desc *d, kcfd_mesh **out) {
    KCFD_ARG(ctx && d && out, "null argument");
    KCFD_ARG(d->n_cells > 0 && d->n_faces > 0, "empty mesh");
    KCFD_ARG(d->ndim == 2 || d->ndim == 3, "ndim must be 2 or 3");
    KCFD_ARG(d->owner && d->neighbour && d->area && d->normal && d->fcenter && d->ccenter &&
                 d->volume && d->d_pn && d->d_pf && d->gweight && d->bpatch && d->bidx &&
                 d->cf_ptr && d->cf_face && d->cf_side && d->indptr && d->indices &&
                 d->slot_on && d->slot_no && d->diag_slot,
             "null mesh array");
    const int n = d->n_cells, nf = d->n_faces;
    // structural checks the kernels rely on
    for (int f = 0; f < nf; ++f) {
        KCFD_ARG(d->owner[f] >= 0 && d->owner[f] < n, "owner out of range");
        KCFD_ARG(d->neighbour[f] < n, "neighbour out of range");
        KCFD_ARG(d->slot_on[f] < d->nnz && d->slot_no[f] < d->nnz, "slot out of range");
        KCFD_ARG(d->neighbour[f] >= 0 || d->bpatch[f] < d->n_patches, "patch out of range");
    }
    KCFD_ARG(d->cf_ptr[0] == 0, "cf_ptr[0] != 0");
    for (int c = 0; c < n; ++c) {
        KCFD_ARG(d->cf_ptr[c + 1] >= d->cf_ptr[c], "cf_ptr not monotone");
        KCFD_ARG(d->diag_slot[c] >= 0 && d->diag_slot[c] < d->nnz, "diag slot out of range");
    }
    const int ncf = d->cf_ptr[n];
    for (int q = 0; q < ncf; ++q) KCFD_ARG(d->cf_face[q] >= 0 && d->cf_face[q] < nf, "cf_face out of range");
    int nbf = 0;
    for (int f = 0; f < nf; ++f) {
        if (d->neighbour[f] < 0) {
            KCFD_ARG(d->bidx[f] >= 0, "boundary face without a boundary index");
            nbf = std::max(nbf, d->bidx[f] + 1);
        }
    }
    // one face per off-diagonal slot (conforming mesh): required for the
    // race-free face-parallel scatter
    {
        std::vector<unsigned char> seen(d->nnz, 0);
        for (int f = 0; f < nf; ++f) {
            if (d->slot_on[f] < 0) continue;
            KCFD_ARG(!seen[d->slot_on[f]] && !seen[d->slot_no[f]],
                     "two faces share one matrix slot (non-conforming mesh)");
            seen[d->slot_on[f]] = seen[d->slot_no[f]] = 1;
        }
    }
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    KCFD_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    kcfd_mesh *m = new kcfd_mesh();
    m->ctx = ctx;
    m->n = n;
    m->nf = nf;
    m->ndim = d->ndim;
    m->npatch = d->n_patches;
    m->nnz = d->nnz;
    m->nbf = nbf;
    m->h_indptr.assign(d->indptr, d->indptr + n + 1);
    m->h_indices.assign(d->indices, d->indices + d->nnz);
    int rc = KCFD_OK;
    auto up = [&](auto **dst, const auto *src, size_t cnt) {
        if (rc == KCFD_OK) rc = upload(dst, src, cnt, s);
    };
    up(&m->owner, d->owner, nf);
    up(&m->nb, d->neighbour, nf);
    up(&m->area, d->area, nf);
    up(&m->normal, d->normal, (size_t)nf * 3);
    up(&m->fc, d->fcenter, (size_t)nf * 3);
    up(&m->cc, d->ccenter, (size_t)n * 3);
    up(&m->vol, d->volume, n);
    up(&m->dpn, d->d_pn, nf);
    up(&m->dpf, d->d_pf, nf);
    up(&m->gw, d->gweight, nf);
    up(&m->bpatch, d->bpatch, nf);
    up(&m->bidx, d->bidx, nf);
    up(&m->cf_ptr, d->cf_ptr, (size_t)n + 1);
    up(&m->cf_face, d->cf_face, ncf);
    up(&m->cf_side, d->cf_side, ncf);
    up(&m->slot_on, d->slot_on, nf);
    up(&m->slot_no, d->slot_no, nf);
    up(&m->diag_slot, d->diag_slot, n);
    up(&m->fco, (const double *)nullptr, nf);
    up(&m->fco2, (const double *)nullptr, nf);
    up(&m->d_kinds, (const int *)nullptr, std::max(1, d->n_patches));
    up(&m->d_vals, (const double *)nullptr, std::max(1, d->n_patches));
    if (rc == KCFD_OK) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            set_error("mesh upload: %s", hipGetErrorString(e));
            rc = KCFD_EHIP;
        }
    }
    if (rc != KCFD_OK) {
        kcfd_mesh_destroy(m);
        return rc;
    }
    *out = m;
    return KCFD_OK;
}

int kcfd_mesh_destroy(kcfd_mesh *m) {
    if (!m) return KCFD_OK;
    std::lock_guard<std::recursive_mutex> g(m->ctx->mu);
    hipSetDevice(m->ctx->device);
    hipStreamSynchronize(m->ctx->stream);
    void *ptrs[] = {m->owner, m->nb, m->area, m->normal, m->fc, m->cc, m->vol, m->dpn,
                    m->dpf, m->gw, m->bpatch, m->bidx, m->cf_ptr, m->cf_face, m->cf_side,
                    m->slot_on, m->slot_no, m->diag_slot, m->fco, m->fco2, m->d_kinds, m->d_vals};
    for (void *p : ptrs)
        if (p) hipFree(p);
    delete m;
    return KCFD_OK;
}

int kcfd_mesh_new_system(kcfd_mesh *m, kcfd_mat **out) {
    KCFD_ARG(m && out, "null argument");
    return kcfd_mat_create(m->ctx, m->n, m->nnz, m->h_indptr.data(), m->h_indices.data(), out);
}

int kcfd_asm_reset(kcfd_mat *a) {
    KCFD_ARG(a != nullptr, "null matrix");
    std::lock_guard<std::recursive_mutex> g(a->ctx->mu);
    ++a->val_gen;
    hipStream_t s = a->ctx->stream;
    if (a->nnz) hipLaunchKernelGGL(k_zero, dim3(blocks(std::min(a->nnz, 1 << 22))), dim3(kBlock), 0, s, a->val, (size_t)a->nnz);
    if (a->n) {
        hipLaunchKernelGGL(k_zero, dim3(blocks(std::min(a->n, 1 << 22))), dim3(kBlock), 0, s, a->rhs, (size_t)a->n);
        hipLaunchKernelGGL(k_zero, dim3(blocks(std::min(a->n, 1 << 22))), dim3(kBlock), 0, s, a->diagt, (size_t)a->n);
    }
    KCFD_HIP(hipGetLastError());
    return KCFD_OK;
}

#define CHECK_PAIR(mesh, mat)                                                              \
    KCFD_ARG((mesh) && (mat), "null argument");                                            \
    KCFD_ARG((mesh)->ctx == (mat)->ctx, "mesh and matrix on different contexts");          \
    KCFD_ARG((mesh)->n == (mat)->n && (mesh)->nnz == (mat)->nnz,                           \
             "matrix does not have the mesh's structure")

int kcfd_asm_diffusion(kcfd_mesh *mesh, kcfd_mat *a, const double *gamma) {
    CHECK_PAIR(mesh, a);
    KCFD_ARG(gamma != nullptr, "null gamma");
    std::lock_guard<std::recursive_mutex> g(a->ctx->mu);
    ++a->val_gen;
    hipStream_t s = a->ctx->stream;
    Timer t(a->ctx);
    MeshView v = view(mesh);
    hipLaunchKernelGGL(k_diff_face, dim3(blocks(mesh->nf)), dim3(kBlock), 0, s, v, gamma, a->val, mesh->fco);
    hipLaunchKernelGGL(k_gather_sym, dim3(blocks(mesh->n)), dim3(kBlock), 0, s, v, mesh->fco, a->val, a->diagt);
    return t.stop();
}

int kcfd_asm_upwind(kcfd_mesh *mesh, kcfd_mat *a, const double *F) {
    CHECK_PAIR(mesh, a);
    KCFD_ARG(F != nullptr, "null flux");
    std::lock_guard<std::recursive_mutex> g(a->ctx->mu);
    ++a->val_gen;
    hipStream_t s = a->ctx->stream;
    Timer t(a->ctx);
    MeshView v = view(mesh);
    hipLaunchKernelGGL(k_upw_face, dim3(blocks(mesh->nf)), dim3(kBlock), 0, s, v, F, a->val);
    hipLaunchKernelGGL(k_upw_cell, dim3(blocks(mesh->n)), dim3(kBlock), 0, s, v, F, a->val, a->diagt);
    return t.stop();
}

int kcfd_asm_boundary(kcfd_mesh *mesh, kcfd_mat *a, const double *gamma, const double *F,
                      const double *phib, const int *kinds, const double *values) {
    CHECK_PAIR(mesh, a);
    KCFD_ARG(gamma && F && kinds && (phib || mesh->nbf == 0), "null argument");
    for (int p = 0; p < mesh->npatch; ++p)
        KCFD_ARG(kinds[p] >= 0 && kinds[p] <= 3, "unknown boundary-condition kind");
    std::lock_guard<std::recursive_mutex> g(a->ctx->mu);
    ++a->val_gen;
    hipStream_t s = a->ctx->stream;
    if (mesh->npatch) {
        std::vector<double> vv(mesh->npatch, 0.0);
        if (values) vv.assign(values, values + mesh->npatch);
        KCFD_HIP(hipMemcpyAsync(mesh->d_kinds, kinds, sizeof(int) * mesh->npatch, hipMemcpyHostToDevice, s));
        KCFD_HIP(hipMemcpyAsync(mesh->d_vals, vv.data(), sizeof(double) * mesh->npatch, hipMemcpyHostToDevice, s));
        KCFD_HIP(hipStreamSynchronize(s));
    }
    Timer t(a->ctx);
    hipLaunchKernelGGL(k_bc_cell, dim3(blocks(mesh->n)), dim3(kBlock), 0, s, view(mesh), gamma, F,
                       phib, mesh->d_kinds, mesh->d_vals, a->val, a->diagt, a->rhs);
    return t.stop();
}

static int cell_terms(kcfd_mesh *mesh, kcfd_mat *a, int mode, const double *a1, const double *a2,
                      double s1, double s2) {
    std::lock_guard<std::recursive_mutex> g(a->ctx->mu);
    if (mode != 1) ++a->val_gen;   // every term but the source writes the diagonal
    hipStream_t s = a->ctx->stream;
    Timer t(a->ctx);
    hipLaunchKernelGGL(k_cell_terms, dim3(blocks(a->n)), dim3(kBlock), 0, s, a->n, mode,
                       mesh ? mesh->diag_slot : a->diagpos, mesh ? mesh->vol : nullptr, a1, a2, s1, s2,
                       a->val, a->diagt, a->rhs);
    return t.stop();
}

int kcfd_asm_temporal(kcfd_mesh *mesh, kcfd_mat *a, double rho, double dt, const double *phi_old) {
    CHECK_PAIR(mesh, a);
    KCFD_ARG(phi_old != nullptr, "null phi_old");
    return cell_terms(mesh, a, 0, phi_old, nullptr, rho, dt);
}

int kcfd_asm_source(kcfd_mesh *mesh, kcfd_mat *a, const double *S) {
    CHECK_PAIR(mesh, a);
    KCFD_ARG(S != nullptr, "null source");
    return cell_terms(mesh, a, 1, S, nullptr, 0.0, 0.0);
}

int kcfd_asm_linearized(kcfd_mesh *mesh, kcfd_mat *a, const double *Sp, const double *Su) {
    CHECK_PAIR(mesh, a);
    KCFD_ARG(Sp && Su, "null source");
    return cell_terms(mesh, a, 2, Sp, Su, 0.0, 0.0);
}

int kcfd_asm_cell_terms(kcfd_mesh *mesh, kcfd_mat *a, const double *kd, const double *ks) {
    CHECK_PAIR(mesh, a);
    KCFD_ARG(kd && ks, "null coefficient");
    return cell_terms(mesh, a, 3, kd, ks, 0.0, 0.0);
}

int kcfd_asm_under_relax(kcfd_mat *a, const double *phi, double alpha) {
    KCFD_ARG(a && phi, "null argument");
    if (alpha >= 1.0) return KCFD_OK;
    KCFD_ARG(alpha > 0.0, "alpha must be positive");
    return cell_terms(nullptr, a, 4, phi, nullptr, alpha, 0.0);
}

int kcfd_asm_pressure(kcfd_mesh *mesh, kcfd_mat *a, const double *dP, double rho,
                      const int *pdirichlet) {
    CHECK_PAIR(mesh, a);
    KCFD_ARG(dP != nullptr, "null dP");
    std::lock_guard<std::recursive_mutex> g(a->ctx->mu);
    ++a->val_gen;
    hipStream_t s = a->ctx->stream;
    if (mesh->npatch) {
        std::vector<int> pd(mesh->npatch, 0);
        if (pdirichlet) pd.assign(pdirichlet, pdirichlet + mesh->npatch);
        KCFD_HIP(hipMemcpyAsync(mesh->d_kinds, pd.data(), sizeof(int) * mesh->npatch, hipMemcpyHostToDevice, s));
        KCFD_HIP(hipStreamSynchronize(s));
    }
    Timer t(a->ctx);
    MeshView v = view(mesh);
    hipLaunchKernelGGL(k_press_face, dim3(blocks(mesh->nf)), dim3(kBlock), 0, s, v, dP, rho, mesh->d_kinds, a->val, mesh->fco);
    hipLaunchKernelGGL(k_press_cell, dim3(blocks(mesh->n)), dim3(kBlock), 0, s, v, mesh->fco, a->val, a->diagt);
    return t.stop();
}

int kcfd_grad_lsq(kcfd_mesh *mesh, const double *phi, const double *phib, double *grad) {
    KCFD_ARG(mesh && phi && grad && (phib || mesh->nbf == 0), "null argument");
    std::lock_guard<std::recursive_mutex> g(mesh->ctx->mu);
    Timer t(mesh->ctx);
    hipLaunchKernelGGL(k_lsq, dim3(blocks(mesh->n)), dim3(kBlock), 0, mesh->ctx->stream, view(mesh), phi, phib,
                       grad);
    return t.stop();
}

int kcfd_simple_dp(kcfd_mesh *mesh, const double *ap0, const double *ap1, const double *ap2, double *dP) {
    KCFD_ARG(mesh && ap0 && ap1 && dP && (mesh->ndim == 2 || ap2), "null argument");
    std::lock_guard<std::recursive_mutex> g(mesh->ctx->mu);
    Timer t(mesh->ctx);
    hipLaunchKernelGGL(k_simple_dp, dim3(blocks(mesh->n)), dim3(kBlock), 0, mesh->ctx->stream, mesh->n,
                       mesh->ndim, mesh->vol, ap0, ap1, mesh->ndim == 3 ? ap2 : ap0, dP);
    return t.stop();
}

int kcfd_simple_rhs(kcfd_mesh *mesh, kcfd_mat *a, const double *mf) {
    CHECK_PAIR(mesh, a);
    KCFD_ARG(mf != nullptr, "null mass flux");
    std::lock_guard<std::recursive_mutex> g(a->ctx->mu);
    Timer t(a->ctx);
    hipLaunchKernelGGL(k_simple_rhs, dim3(blocks(mesh->n)), dim3(kBlock), 0, a->ctx->stream, view(mesh), mf,
                       a->rhs);
    return t.stop();
}

int kcfd_simple_fix_reference(kcfd_mat *a) {
    KCFD_ARG(a != nullptr, "null matrix");
    KCFD_ARG(a->n > 0 && !a->h_diagpos.empty() && a->h_diagpos[0] >= 0, "row 0 has no diagonal entry");
    std::lock_guard<std::recursive_mutex> g(a->ctx->mu);
    const int len = a->h_indptr[1] - a->h_indptr[0];
    KCFD_ARG(len <= 1024, "row 0 longer than 1024 entries");
    ++a->val_gen;
    Timer t(a->ctx);
    hipLaunchKernelGGL(k_simple_fix_row0, dim3(1), dim3(1024), 0, a->ctx->stream, a->indptr, a->diagpos, a->val,
                       a->rhs);
    return t.stop();
}

int kcfd_simple_correct(kcfd_mesh *mesh, const double *pp, const double *dP, const int *pdirichlet,
                        double alpha_p, double *U, double *p) {
    KCFD_ARG(mesh && pp && dP && U && p, "null argument");
    std::lock_guard<std::recursive_mutex> g(mesh->ctx->mu);
    hipStream_t s = mesh->ctx->stream;
    if (mesh->npatch) {
        std::vector<int> pd(mesh->npatch, 0);
        if (pdirichlet) pd.assign(pdirichlet, pdirichlet + mesh->npatch);
        KCFD_HIP(hipMemcpyAsync(mesh->d_kinds, pd.data(), sizeof(int) * mesh->npatch, hipMemcpyHostToDevice, s));
        KCFD_HIP(hipStreamSynchronize(s));
    }
    Timer t(mesh->ctx);
    hipLaunchKernelGGL(k_simple_correct, dim3(blocks(mesh->n)), dim3(kBlock), 0, s, view(mesh), pp, dP,
                       mesh->d_kinds, alpha_p, U, p);
    return t.stop();
}

int kcfd_twofluid_pressure(kcfd_mesh *mesh, kcfd_mat *a, const double *alpha_l, const double *alpha_g,
                           const double *aP_l, const double *aP_g, double rho_l, double rho_g,
                           const double *mf_l, const double *mf_g) {
    CHECK_PAIR(mesh, a);
    KCFD_ARG(alpha_l && alpha_g && aP_l && aP_g && mf_l && mf_g, "null argument");
    KCFD_ARG(a->n > 0 && !a->h_diagpos.empty() && a->h_diagpos[0] >= 0, "row 0 has no diagonal entry");
    std::lock_guard<std::recursive_mutex> g(a->ctx->mu);
    ++a->val_gen;
    hipStream_t s = a->ctx->stream;
    Timer t(a->ctx);
    MeshView v = view(mesh);
    hipLaunchKernelGGL(k_tf_press_face, dim3(blocks(mesh->nf)), dim3(kBlock), 0, s, v, alpha_l, alpha_g, aP_l,
                       aP_g, rho_l, rho_g, mf_l, mf_g, a->val, mesh->fco, mesh->fco2);
    hipLaunchKernelGGL(k_press_cell, dim3(blocks(mesh->n)), dim3(kBlock), 0, s, v, mesh->fco, a->val, a->diagt);
    hipLaunchKernelGGL(k_simple_rhs, dim3(blocks(mesh->n)), dim3(kBlock), 0, s, v, mesh->fco2, a->rhs);
    hipLaunchKernelGGL(k_tf_ref_cell, dim3(1), dim3(64), 0, s, a->diagpos, a->val, a->diagt);
    return t.stop();
}

int kcfd_twofluid_correct(kcfd_mesh *mesh, const double *pp, const double *alpha_l, const double *alpha_g,
                          const double *aP_l, const double *aP_g, double alpha_p, double *U_l, double *U_g,
                          double *p) {
    KCFD_ARG(mesh && pp && alpha_l && alpha_g && aP_l && aP_g && U_l && U_g && p, "null argument");
    std::lock_guard<std::recursive_mutex> g(mesh->ctx->mu);
    Timer t(mesh->ctx);
    hipLaunchKernelGGL(k_tf_correct, dim3(blocks(mesh->n)), dim3(kBlock), 0, mesh->ctx->stream, view(mesh), pp,
                       alpha_l, alpha_g, aP_l, aP_g, alpha_p, U_l, U_g, p);
    return t.stop();
}

int kcfd_face_values(kcfd_mesh *mesh, int mode, const double *phi, const double *phib, const double *F,
                     const double *grad, double *phif) {
    KCFD_ARG(mesh && phi && phif && (phib || mesh->nbf == 0), "null argument");
    KCFD_ARG(mode >= 0 && mode <= 2, "mode must be 0 (upwind), 1 (central) or 2 (TVD)");
    KCFD_ARG(mode == 1 || F, "upwind and TVD values need the face fluxes");
    KCFD_ARG(mode != 2 || grad, "TVD values need the gradient");
    std::lock_guard<std::recursive_mutex> g(mesh->ctx->mu);
    Timer t(mesh->ctx);
    hipLaunchKernelGGL(k_face_values, dim3(blocks(mesh->nf)), dim3(kBlock), 0, mesh->ctx->stream, view(mesh),
                       mode, phi, phib, F, grad, phif);
    return t.stop();
}

int kcfd_schiller_naumann(kcfd_ctx *ctx, const double *re, double *cd, long long n) {
    KCFD_ARG(ctx && n >= 0 && (n == 0 || (re && cd)), "null argument");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (n == 0) return KCFD_OK;
    hipLaunchKernelGGL(k_schiller_naumann, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       ctx->stream, re, cd, (size_t)n);
    KCFD_HIP(hipGetLastError());
    KCFD_HIP(hipStreamSynchronize(ctx->stream));
    return KCFD_OK;
}

int kcfd_drag_implicit(kcfd_ctx *ctx, int n, int ndim, const double *alpha_g, double rho_l,
                       const double *u_g, const double *u_l, double d_b, double mu_l, double *K) {
    KCFD_ARG(ctx && n >= 0 && (n == 0 || (alpha_g && u_g && u_l && K)), "null argument");
    KCFD_ARG(ndim >= 1 && ndim <= 3, "ndim must be 1..3");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (n == 0) return KCFD_OK;
    Timer t(ctx);
    hipLaunchKernelGGL(k_drag, dim3(blocks(n)), dim3(kBlock), 0, ctx->stream, n, ndim, alpha_g, rho_l, u_g, u_l,
                       d_b, mu_l, K);
    return t.stop();
}

int kcfd_limiter(kcfd_ctx *ctx, int kind, const double *r, double *psi, long long n) {
    KCFD_ARG(ctx && n >= 0 && (n == 0 || (r && psi)), "null argument");
    KCFD_ARG(kind >= KCFD_LIM_MINMOD && kind <= KCFD_LIM_VAN_LEER, "unknown limiter");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (n == 0) return KCFD_OK;
    hipLaunchKernelGGL(k_limiter, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, kind,
                       r, psi, (size_t)n);
    KCFD_HIP(hipGetLastError());
    KCFD_HIP(hipStreamSynchronize(ctx->stream));
    return KCFD_OK;
}

int kcfd_grad_gg(kcfd_mesh *mesh, const double *phi, const double *phib, double *grad) {
    KCFD_ARG(mesh && phi && grad && (phib || mesh->nbf == 0), "null argument");
    std::lock_guard<std::recursive_mutex> g(mesh->ctx->mu);
    Timer t(mesh->ctx);
    hipLaunchKernelGGL(k_gg, dim3(blocks(mesh->n)), dim3(kBlock), 0, mesh->ctx->stream, view(mesh), phi, phib, grad);
    return t.stop();
}

int kcfd_muscl(kcfd_mesh *mesh, const double *phi, const double *F, const double *grad, int lim,
               double *corr) {
    KCFD_ARG(mesh && phi && F && grad && corr, "null argument");
    KCFD_ARG(lim >= 0 && lim <= 3, "unknown limiter");
    std::lock_guard<std::recursive_mutex> g(mesh->ctx->mu);
    hipStream_t s = mesh->ctx->stream;
    Timer t(mesh->ctx);
    MeshView v = view(mesh);
    hipLaunchKernelGGL(k_muscl_face, dim3(blocks(mesh->nf)), dim3(kBlock), 0, s, v, phi, F, grad, lim, mesh->fco, mesh->fco2);
    hipLaunchKernelGGL(k_muscl_cell, dim3(blocks(mesh->n)), dim3(kBlock), 0, s, v, mesh->fco, mesh->fco2, corr);
    return t.stop();
}

int kcfd_mass_flux(kcfd_mesh *mesh, const double *U, const double *Ub, double rho, double *F) {
    KCFD_ARG(mesh && U && F && (Ub || mesh->nbf == 0), "null argument");
    std::lock_guard<std::recursive_mutex> g(mesh->ctx->mu);
    Timer t(mesh->ctx);
    hipLaunchKernelGGL(k_mflux, dim3(blocks(mesh->nf)), dim3(kBlock), 0, mesh->ctx->stream, view(mesh), U, Ub, rho, F);
    return t.stop();
}

int kcfd_last_kernel_ms(kcfd_ctx *ctx, double *ms) {
    KCFD_ARG(ctx && ms, "null argument");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    KCFD_HIP(hipEventSynchronize(ctx->ev1));
    float t = 0.f;
    KCFD_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    *ms = t;
    return KCFD_OK;
}

}  // extern "C"
