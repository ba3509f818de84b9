// Flow solver around the GPU coefficient assembly.
//
// Assembly kernels: one thread per cell gathers its row of the system from
// the cell's faces (ascending face index, then its boundary faces, then the
// cell terms) and writes the whole row of the caller's CSR value array, the
// right-hand side and the initial guess.  Every coefficient of a face is
// computed by both cells that share it, so there are no atomics and the sums
// run in the reference's order (the sequential face loop, then the boundary
// loop, then the cell loop: cflow_solver.cpp:268-283, 323-342, 429-446), which
// keeps the assembly deterministic.
//
// Support loops: the rest of CFlow_solver::iterate (boundary updates, face
// mass fluxes, gradients, corrections, field updates: cflow_solver.cpp:229-476,
// cflow_solver_supply.cpp) as device kernels over the resident state (the
// default), or restated in C++ on host arrays that are uploaded for each
// equation (OCSL_SUPPORT_HOST, a cross-check).  The gradient kernels gather
// per cell in the same face order as the reference's scatter loops, so each
// cell's sum is added in the reference's order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "ocsl_flow.h"
#include "ocsl_handle.h"
#include "ocsl_kernels.h"
#include "ocsl_order.h"

namespace ocsl {

constexpr int kMaxRow = 16;           // stored entries per row (sentinels included)
constexpr double kAlpha = 0.7;        // blended-central weight (solvers_common.h:30)
constexpr double kUpwind = 1.0 - kAlpha;
constexpr double kRelaxP = 0.5;       // explicit_underrelaxation_P (cflow_solver.h:178)
constexpr double kRelaxU = 0.8;       // explicit_underrelaxation_u (cflow_solver.h:179)

struct BcDev {
    int type, no_slip, pad0, pad1;
};

// Gas model (cfluid_properties.cpp), fixed heat capacity.
struct Gas {
    double R, gamma, cp, t_ref, h_f, s0, s1, pr;
    __host__ __device__ double density(double p, double T) const { return p / R / T; }
    __host__ __device__ double dp_drho(double T) const {
        double r = R * T;
        r = r * gamma;
        return 1.0 / r;
    }
    __host__ __device__ double sound(double T) const {
        double r = R * T;
        return sqrt(r * gamma);
    }
    __host__ __device__ double viscosity(double T) const {
        double r = s0 * pow(T, 3.0 / 2.0);
        return r / (T + s1);
    }
    __host__ __device__ double enthalpy(double T, const double* v) const {
        const double vm = mag3(v);
        double r = h_f;
        r += (T - t_ref) * cp;
        r += vm * vm * 0.5;
        return r;
    }
    __host__ __device__ double temperature(double h, const double* v) const {
        const double vm = mag3(v);
        double r = t_ref + (h - h_f - vm * vm * 0.5) / cp;
        if (r <= 200.0) r = 200.0;
        if (r >= 6000.0) r = 6000.0;
        return r;
    }
    __host__ __device__ static double dot3(const double* a, const double* b) {
        double r = 0.0;
        for (int i = 0; i < 3; ++i) r += a[i] * b[i];
        return r;
    }
    __host__ __device__ static double mag3(const double* a) { return sqrt(dot3(a, a)); }
};

struct FlowPhys {
    Gas gas;
    double urf;  // implicit under-relaxation (steady)
    int n_bc;
    BcDev bc[OCSL_FLOW_MAX_BC];
};

struct MeshDev {
    int n, F, B;
    const double* vol;
    const int *fL, *fR;
    const double *fS, *fe, *fdist, *fksi, *fgeo, *falpha;
    const int *bL, *bmark;
    const double *bS, *barea, *bksi, *bperp, *be, *balpha, *bdist;
    // Per cell, its interior faces in ascending caller face order, sliced 64
    // cells (one wave) to a slice: entry k of cell i is at cf_sp[i / 64] +
    // 64 k + i % 64, so each step of the face loop is one contiguous load per
    // wave; code = f << 5 | pos << 1 | side (pos: the neighbour's slot
    // position in the row), or -1 after the cell's last face.
    const int *cf_sp, *cf_code;
    const int *cb_ptr, *cb_face;           // per cell: boundary faces
    const int *row_ptr, *dpos;             // caller layout (int32), diagonal position in the row
    // per cell-face entry, beside cf_code: the cell across the face
    const int* cf_nb;
    // Cells stored in an order of their own (cperm != nullptr): device cell
    // k is the caller's cell cperm[k], whose row starts at rstart[k] in the
    // caller's value array and holds rlen[k] slots.
    const int *cperm, *rstart, *rlen;
};

struct V3 {
    double v[3];
};

// The 3-vectors of cells (u, u_old, gp, gpc, ap_vec) and of interior faces
// (f_stress) are stored as three planes, component k of item i at
// k * stride + i (cs cells, fs faces): a momentum component's kernels read
// one plane, not every third element of interleaved triples.  Boundary-face
// vectors (b_u) stay interleaved.
struct StateDev {
    double *u, *u_old, *rho, *rho_old, *p, *p_old, *pc, *gp, *T, *h, *h_old, *ap, *ap_vec, *dt;
    double *f_mdot, *f_rho, *f_qy, *f_stress;
    double *b_T, *b_rho, *b_mdot, *b_h, *b_qy, *b_u;
    int* b_super;
    double *gpc, *mach, *b_p, *b_p_out;
    int cs, fs;  // plane strides: cells, interior faces
    __device__ __forceinline__ V3 vec3(const double* a, int i) const {
        return {{a[i], a[cs + i], a[2 * cs + i]}};
    }
};

enum { kMomentum = 0, kPressure = 1, kEnthalpy = 2 };

// A cell's row staged in LDS: slot q at p[q * stride] (stride 1: the row
// laid out as in the caller's value array; the block size: one column of
// slots per thread, free of bank conflicts).
struct RowAcc {
    double* p;
    int stride;
    __device__ __forceinline__ double& operator[](int q) const { return p[q * stride]; }
};

// a: the cell's row; pos: the slot position in the row.  The reference adds
// each term, already rounded, to the slot: the empty asm keeps the compiler
// from fusing the product that formed v into this add.
__device__ __forceinline__ void add_at(RowAcc a, int pos, double v) {
    asm volatile("" : "+v"(v));
    a[pos] += v;
}

__device__ __forceinline__ double face_avg(double geo, double l, double r) {
    double v = geo * l;
    v += (1.0 - geo) * r;
    return v;
}

template <bool kBlend>
__device__ __forceinline__ double conv_interp(double geo, double mdot, double l, double r) {
    if (mdot >= 0.0) return kBlend ? l * kUpwind + kAlpha * face_avg(geo, l, r) : l;
    return kBlend ? r * kUpwind + kAlpha * face_avg(geo, l, r) : r;
}

// Convective_flux (solvers_common.h:90-135) seen from one cell of the face:
// side 0 = left (owner), 1 = right; dp / op: positions of the diagonal and of
// the other cell's column in this row.
template <bool kBlend>
__device__ __forceinline__ void conv_row(RowAcc a, double& rhs, int side, int dp, int op,
                                         double mdot, double l, double r, double geo) {
    if (mdot >= 0.0) {
        if (side == 0) add_at(a, dp, mdot);
        else add_at(a, op, -mdot);
        if constexpr (kBlend) {
            const double val = kAlpha * (face_avg(geo, l, r) - l);
            rhs += side == 0 ? -mdot * val : mdot * val;
        }
    } else if (mdot < 0.0) {
        if (side == 0) add_at(a, op, mdot);
        else add_at(a, dp, -mdot);
        if constexpr (kBlend) {
            const double val = kAlpha * (face_avg(geo, l, r) - r);
            rhs += side == 0 ? -mdot * val : mdot * val;
        }
    }
}

// Diffusive_flux (solvers_common.h:42-53): +D on the diagonal, -D off it.
__device__ __forceinline__ void diff_row(RowAcc a, int dp, int op, double D) {
    add_at(a, dp, D);
    add_at(a, op, -D);
}

template <int kEq, bool kVisc, bool kBlend, bool kTransient>
__device__ __forceinline__ void assemble_cell(const MeshDev& m, const StateDev& s, const FlowPhys& ph, int d,
                                              int i, int oi, RowAcc a, double* __restrict__ b,
                                              double* __restrict__ x0) {
    const Gas& gas = ph.gas;
    double rhs = 0.0;
    const int dp = m.dpos[i];
    // this cell's own values, the same for each of its faces
    double own_q = 0.0, own_T = 0.0;
    if constexpr (kEq == kMomentum) own_q = s.u[d * s.cs + i];
    else if constexpr (kEq == kPressure) own_q = s.pc[i];
    else own_q = s.h[i];
    if constexpr (kEq != kMomentum || kVisc) own_T = s.T[i];
    // interior faces (cflow_solver_cell_face_loops.cpp:38-58, 186-218, 334-346);
    // (l, r) are the face's left and right cell values whichever side this is
    for (int e = m.cf_sp[i >> 6] + (i & 63), e1 = m.cf_sp[(i >> 6) + 1]; e < e1; e += 64) {
        const int code = m.cf_code[e];
        if (code < 0) break;
        const int f = code >> 5, side = code & 1, op = (code >> 1) & (kMaxRow - 1);
        const int j = m.cf_nb[e];
        const double geo = m.fgeo[f];
        const double mdot = s.f_mdot[f];
        if constexpr (kEq == kMomentum) {
            const double qj = s.u[d * s.cs + j];
            conv_row<kBlend>(a, rhs, side, dp, op, mdot, side ? qj : own_q, side ? own_q : qj, geo);
            if constexpr (kVisc) {
                const double Tj = s.T[j];
                const double Tf = face_avg(geo, side ? Tj : own_T, side ? own_T : Tj);
                double D = gas.viscosity(Tf);
                D = D * m.fksi[f];
                diff_row(a, dp, op, D);
                const double st = s.f_stress[d * s.fs + f];
                rhs += side == 0 ? st : -st;
            }
        } else if constexpr (kEq == kPressure) {
            rhs += side == 0 ? -mdot : mdot;
            const double Tj = s.T[j], qj = s.pc[j];
            const double Tf = conv_interp<kBlend>(geo, mdot, side ? Tj : own_T, side ? own_T : Tj);
            double conv_w = gas.dp_drho(Tf);
            conv_w = conv_w / s.f_rho[f];
            conv_w *= mdot;
            conv_row<kBlend>(a, rhs, side, dp, op, conv_w, side ? qj : own_q, side ? own_q : qj, geo);
            const double diff_w = s.f_qy[f] * s.f_rho[f];
            diff_row(a, dp, op, diff_w);
        } else {
            const double Tj = s.T[j], qj = s.h[j];
            conv_row<kBlend>(a, rhs, side, dp, op, mdot, side ? qj : own_q, side ? own_q : qj, geo);
            const double Tf = face_avg(geo, side ? Tj : own_T, side ? own_T : Tj);
            double diff_w = gas.viscosity(Tf);
            diff_w = diff_w / gas.pr * m.fksi[f];
            diff_row(a, dp, op, diff_w);
        }
    }
    // boundary faces (cflow_solver_cell_face_loops.cpp:77-132, 229-297, 357-421)
    for (int e = m.cb_ptr[i]; e < m.cb_ptr[i + 1]; ++e) {
        const int bf = m.cb_face[e];
        const BcDev bc = ph.bc[m.bmark[bf]];
        const double bm = s.b_mdot[bf];
        if constexpr (kEq == kMomentum) {
            if (bc.type == OCSL_BC_WALL) {
                if constexpr (kVisc) {
                    if (bc.no_slip) {
                        const double inv = 1.0 / m.barea[bf];
                        double nrm[3];
                        for (int k = 0; k < 3; ++k) nrm[k] = inv * m.bS[3 * bf + k];
                        double D = gas.viscosity(s.T[i]);
                        D *= m.barea[bf] / m.bperp[bf];
                        const double a_wall = D * (1.0 - (nrm[d] * nrm[d]));
                        add_at(a, dp, a_wall);
                        double tang = 0.0;
                        for (int k = d + 1; k < d + 3; ++k) {
                            const int kk = k % 3;
                            tang += s.u[kk * s.cs + i] * nrm[kk] * nrm[d];
                        }
                        tang = D * tang;
                        rhs += tang;
                    }
                }
            }
            if (bc.type == OCSL_BC_VELOCITY_INLET || bc.type == OCSL_BC_SUPERSONIC_INLET) {
                const double v = bm * s.b_u[3 * bf + d];
                rhs += -v;
            }
            if (bc.type == OCSL_BC_PRESSURE_OUTLET) add_at(a, dp, bm);
        } else if constexpr (kEq == kPressure) {
            if (bc.type == OCSL_BC_VELOCITY_INLET) {
                rhs += -bm;
                double term = gas.dp_drho(s.b_T[bf]) / s.rho[i];
                term *= bm;
                if (!isnan(term)) add_at(a, dp, term);  // Convective_flux_neumann_zero_BC adds any non-NaN flux
            }
            if (bc.type == OCSL_BC_SUPERSONIC_INLET) {
                rhs += -bm;
                const double diff_w = s.b_qy[bf] * s.b_rho[bf];
                add_at(a, dp, diff_w);
                rhs += diff_w * 0.0;
            }
            if (bc.type == OCSL_BC_PRESSURE_OUTLET) {
                rhs += -bm;
                if (s.b_super[bf]) {
                    double term = gas.dp_drho(s.T[i]) / s.rho[i];
                    term *= bm;
                    if (!isnan(term)) add_at(a, dp, term);
                } else {
                    const double diff_w = s.b_qy[bf] * s.rho[i];
                    add_at(a, dp, diff_w);
                    rhs += diff_w * 0.0;
                }
            }
        } else {
            if (bc.type == OCSL_BC_VELOCITY_INLET || bc.type == OCSL_BC_SUPERSONIC_INLET) {
                const double v = bm * s.b_h[bf];
                rhs += -v;
                double diff_w = gas.viscosity(s.b_T[bf]);
                diff_w *= gas.pr * m.bksi[bf];  // mu * Pr: the reference's boundary form
                add_at(a, dp, diff_w);
                rhs += diff_w * s.b_h[bf];
            }
            if (bc.type == OCSL_BC_PRESSURE_OUTLET) {
                if (!isnan(bm)) add_at(a, dp, bm);
            }
        }
    }
    // cell terms (cflow_solver_cell_face_loops.cpp:143-175, 299-322, 423-446)
    const double V = m.vol[i];
    double diag = a[dp];
    if constexpr (kEq == kMomentum) {
        double term = 0.0, term2 = 0.0;
        if constexpr (kTransient) {
            term = V / s.dt[i];
            term2 = term * s.rho[i];
            diag += term2;
            term2 = term * (s.u_old[d * s.cs + i] * s.rho_old[i]);
        }
        term2 = term2 - s.gp[d * s.cs + i] * V;
        rhs += term2;
        if constexpr (!kTransient) {
            const double dd = diag;
            diag = dd / ph.urf;
            rhs += (1.0 - ph.urf) / ph.urf * dd * s.u_old[d * s.cs + i];
        }
        s.ap[i] += diag / 3.0;
        s.ap_vec[d * s.cs + i] = diag;
        x0[oi] = s.u[d * s.cs + i];
    } else if constexpr (kEq == kPressure) {
        if constexpr (kTransient) {
            const double term = V / s.dt[i];
            double term2 = term * (s.rho[i] - s.rho_old[i]);
            rhs += -term2;
            term2 = term * gas.dp_drho(s.T[i]);
            diag += term2;
        } else {
            diag = diag / ph.urf;
        }
        x0[oi] = s.pc[i];
    } else {
        if constexpr (kTransient) {
            const double term = V / s.dt[i];
            double term2 = term * s.rho[i];
            diag += term2;
            term2 = term * (s.h_old[i] * s.rho_old[i]);
            term2 = term2 + (s.p[i] - s.p_old[i]) / s.dt[i] * V;
            rhs += term2;
        } else {
            const double dd = diag;
            diag = dd / ph.urf;
            rhs += (1.0 - ph.urf) / ph.urf * dd * s.h_old[i];
        }
        x0[oi] = s.h[i];
    }
    a[dp] = diag;
    b[oi] = rhs;
}

// One thread per cell; each XCD takes a contiguous range of cells.  Rows are
// accumulated in LDS.  Cells in the caller's order (kPerm false): the block's
// rows lie together in the caller's value array, so they are staged as laid
// out there (sentinel slots stay 0) and written out as one coalesced run.
// Cells in an order of their own (kPerm): each thread's row sits in a column
// of LDS slots; the block then writes the rows to the caller's rows of its
// cells, eight lanes to a row; b and x0 go to the caller's index of the cell.
template <int kEq, bool kVisc, bool kBlend, bool kTransient, bool kPerm>
__global__ void __launch_bounds__(kBlock) k_assemble(MeshDev m, StateDev s, FlowPhys ph, int d,
                                                     double* __restrict__ val, double* __restrict__ b,
                                                     double* __restrict__ x0) {
    extern __shared__ double rows[];
    const int i0 = xcd_block(blockIdx.x, gridDim.x) * blockDim.x, i = i0 + threadIdx.x;
    if constexpr (!kPerm) {
        const int base = m.row_ptr[i0], span = m.row_ptr[min(i0 + (int)blockDim.x, m.n)] - base;
        for (int k = threadIdx.x; k < span; k += blockDim.x) rows[k] = 0.0;
        __syncthreads();
        if (i < m.n)
            assemble_cell<kEq, kVisc, kBlend, kTransient>(m, s, ph, d, i, i, RowAcc{rows + (m.row_ptr[i] - base), 1},
                                                          b, x0);
        __syncthreads();
        for (int k = threadIdx.x; k < span; k += blockDim.x) val[base + k] = rows[k];
    } else {
        // slot q of thread t's row at rows[q * (nt + 1) + t]: conflict-free
        // both for the per-thread accumulation and for the write-out below
        __shared__ int s_rs[kBlock], s_len[kBlock];
        const int t = threadIdx.x, nt = blockDim.x, ld = nt + 1;
        int len = 0, rs = 0;
        if (i < m.n) {
            len = m.rlen[i];
            rs = m.rstart[i];
            for (int q = 0; q < len; ++q) rows[q * ld + t] = 0.0;
            assemble_cell<kEq, kVisc, kBlend, kTransient>(m, s, ph, d, i, m.cperm[i], RowAcc{rows + t, ld}, b, x0);
        }
        s_rs[t] = rs;
        s_len[t] = len;
        __syncthreads();
        // eight lanes per row: each store writes eight rows' contiguous runs
        // instead of one slot of 64 scattered rows
        for (int r = t >> 3; r < nt; r += nt >> 3) {
            const int L = s_len[r], base = s_rs[r];
            for (int q = t & 7; q < L; q += 8) val[base + q] = rows[q * ld + r];
        }
    }
}

// calc_flux_explicit_visc_part (cflow_solver_cell_face_loops.cpp:25-35):
// mu(T_f) ((grad u)^T - 2/3 (div u) I) S_f with the two-point gradient
// (tensor.h:194-227).
__global__ void __launch_bounds__(kBlock) k_visc_stress(MeshDev m, StateDev s, Gas gas) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const int L = m.fL[f], R = m.fR[f];
    const double dist = m.fdist[f];
    double du[3], e[3];
    for (int k = 0; k < 3; ++k) {
        du[k] = s.u[k * s.cs + R] - s.u[k * s.cs + L];
        e[k] = m.fe[3 * f + k];
    }
    double M[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[r][c] = du[c] * e[r] / dist;  // transpose of du_i e_j / d
    double divd = 0.0;
    for (int k = 0; k < 3; ++k) divd += du[k] * e[k] / dist;
    const double two3 = 2.0 / 3.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[r][c] = M[r][c] - two3 * (r == c ? divd : 0.0);
    const double mu = gas.viscosity(face_avg(m.fgeo[f], s.T[L], s.T[R]));
    for (int r = 0; r < 3; ++r) {
        double acc = 0.0;
        for (int c = 0; c < 3; ++c) acc += (mu * M[r][c]) * m.fS[3 * f + c];
        s.f_stress[r * s.fs + f] = acc;
    }
}


// ---------------------------------------------------------------- support loops
// Face interpolation of calc_face_mdot / correct_face_mdot: the convective
// interpolation (upwind or blended) by the sign of the mass flux.
__device__ __forceinline__ double conv_interp_d(int blend, double geo, double mdot, double l, double r) {
    return blend ? conv_interp<true>(geo, mdot, l, r) : conv_interp<false>(geo, mdot, l, r);
}

// set_boundary_conditions (cflow_solver.cpp:192-227), one thread per boundary face.
__global__ void __launch_bounds__(kBlock) k_set_bcs(MeshDev m, StateDev s, FlowPhys ph) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.B) return;
    const Gas& g = ph.gas;
    const int L = m.bL[b], t = ph.bc[m.bmark[b]].type;
    if (t == OCSL_BC_VELOCITY_INLET) {
        s.b_p[b] = s.p[L];
        s.b_rho[b] = g.density(s.p[L], s.b_T[b]);
        s.b_mdot[b] = s.b_rho[b] * Gas::dot3(s.b_u + 3 * b, m.bS + 3 * b);
        s.b_h[b] = g.enthalpy(s.b_T[b], s.b_u + 3 * b);
    }
    if (t == OCSL_BC_PRESSURE_OUTLET) {
        s.b_T[b] = s.T[L];
        for (int k = 0; k < 3; ++k) s.b_u[3 * b + k] = s.u[k * s.cs + L];
        s.b_rho[b] = g.density(s.b_p[b], s.T[L]);
        s.b_h[b] = g.enthalpy(s.T[L], s.vec3(s.u, L).v);
        // both branches mark the face supersonic (cflow_solver.cpp:208-217)
        s.b_super[b] = 1;
        if ((s.b_mdot[b] / s.rho[L] / m.barea[b]) >= g.sound(s.T[L])) s.b_p[b] = s.p[L];
        else s.b_p[b] = s.b_p_out[b];
    }
    if (t == OCSL_BC_WALL) s.b_p[b] = s.p[L] + Gas::dot3(s.vec3(s.gp, L).v, m.be + 3 * b) * m.bdist[b];
}

// calc_press_grad (cflow_solver_supply.cpp:132-175) gathered per cell: the
// cell's faces in ascending order (+ as left cell, - as right), then its
// boundary faces, then the division by the volume.
__global__ void __launch_bounds__(kBlock) k_press_grad(MeshDev m, StateDev s, FlowPhys ph) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    double gv[3] = {0.0, 0.0, 0.0};
    const double ap_i = s.ap[i], p_i = s.p[i];
    for (int e = m.cf_sp[i >> 6] + (i & 63), e1 = m.cf_sp[(i >> 6) + 1]; e < e1; e += 64) {
        const int code = m.cf_code[e];
        if (code < 0) break;
        const int f = code >> 5, j = m.cf_nb[e];
        const double ap_j = s.ap[j], p_j = s.p[j];
        const double apL = code & 1 ? ap_j : ap_i, apR = code & 1 ? ap_i : ap_j;
        const double pL = code & 1 ? p_j : p_i, pR = code & 1 ? p_i : p_j;
        double fp = (apL + apR);
        fp = (apL * pL + apR * pR) / fp;
        for (int k = 0; k < 3; ++k) {
            if (code & 1) gv[k] = gv[k] - fp * m.fS[3 * f + k];
            else gv[k] = gv[k] + fp * m.fS[3 * f + k];
        }
    }
    for (int e = m.cb_ptr[i]; e < m.cb_ptr[i + 1]; ++e) {
        const int b = m.cb_face[e], t = ph.bc[m.bmark[b]].type;
        double pv;
        if (t == OCSL_BC_WALL || t == OCSL_BC_SYMMETRY || t == OCSL_BC_EMPTY || t == OCSL_BC_VELOCITY_INLET)
            pv = s.p[i];
        else if (t == OCSL_BC_PRESSURE_OUTLET || t == OCSL_BC_SUPERSONIC_INLET)
            pv = s.b_p[b];
        else
            continue;
        for (int k = 0; k < 3; ++k) gv[k] = gv[k] + pv * m.bS[3 * b + k];
    }
    for (int k = 0; k < 3; ++k) s.gp[k * s.cs + i] = (1.0 / m.vol[i]) * gv[k];
}

// calc_press_corr_grad (cflow_solver_supply.cpp:177-214), gathered per cell.
__global__ void __launch_bounds__(kBlock) k_press_corr_grad(MeshDev m, StateDev s, FlowPhys ph) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    double gv[3] = {0.0, 0.0, 0.0};
    const double pc_i = s.pc[i];
    for (int e = m.cf_sp[i >> 6] + (i & 63), e1 = m.cf_sp[(i >> 6) + 1]; e < e1; e += 64) {
        const int code = m.cf_code[e];
        if (code < 0) break;
        const int f = code >> 5;
        const double geo = m.fgeo[f], pc_j = s.pc[m.cf_nb[e]];
        double fp = geo * (code & 1 ? pc_j : pc_i);
        fp += (1.0 - geo) * (code & 1 ? pc_i : pc_j);
        for (int k = 0; k < 3; ++k) {
            if (code & 1) gv[k] = gv[k] - fp * m.fS[3 * f + k];
            else gv[k] = gv[k] + fp * m.fS[3 * f + k];
        }
    }
    for (int e = m.cb_ptr[i]; e < m.cb_ptr[i + 1]; ++e) {
        const int b = m.cb_face[e], t = ph.bc[m.bmark[b]].type;
        if (t == OCSL_BC_WALL || t == OCSL_BC_SYMMETRY || t == OCSL_BC_EMPTY || t == OCSL_BC_VELOCITY_INLET ||
            (s.b_super[b] && t == OCSL_BC_PRESSURE_OUTLET))
            for (int k = 0; k < 3; ++k) gv[k] = gv[k] + s.pc[i] * m.bS[3 * b + k];
    }
    for (int k = 0; k < 3; ++k) s.gpc[k * s.cs + i] = (1.0 / m.vol[i]) * gv[k];
}

// calc_face_mdot, interior faces (cflow_solver_supply.cpp:25-46).
__global__ void __launch_bounds__(kBlock) k_face_mdot(MeshDev m, StateDev s, int blend) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const int L = m.fL[f], R = m.fR[f];
    const double geo = m.fgeo[f];
    const double qy = ((m.vol[L] / s.ap[L]) + (m.vol[R] / s.ap[R])) * Gas::dot3(m.falpha + 3 * f, m.fS + 3 * f);
    s.f_qy[f] = qy;
    double pgf[3], vf[3], dvec[3];
    for (int k = 0; k < 3; ++k) {
        pgf[k] = geo * s.gp[k * s.cs + L];
        pgf[k] += (1.0 - geo) * s.gp[k * s.cs + R];
        vf[k] = geo * s.u[k * s.cs + L];
        vf[k] += (1.0 - geo) * s.u[k * s.cs + R];
        dvec[k] = m.fdist[f] * m.fe[3 * f + k];
    }
    const double YF = qy * (s.p[R] - s.p[L] - Gas::dot3(pgf, dvec));
    double mdot = Gas::dot3(vf, m.fS + 3 * f) - YF;
    const double rf = conv_interp_d(blend, geo, mdot, s.rho[L], s.rho[R]);
    s.f_rho[f] = rf;
    mdot *= rf;
    s.f_mdot[f] = mdot;
}

// calc_face_mdot, boundary faces (cflow_solver_supply.cpp:54-93).
__global__ void __launch_bounds__(kBlock) k_bc_mdot(MeshDev m, StateDev s, FlowPhys ph) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.B) return;
    const Gas& g = ph.gas;
    const int L = m.bL[b], t = ph.bc[m.bmark[b]].type;
    if (t == OCSL_BC_VELOCITY_INLET)
        s.b_mdot[b] = g.density(s.p[L], s.b_T[b]) * Gas::dot3(s.b_u + 3 * b, m.bS + 3 * b);
    if (t == OCSL_BC_SUPERSONIC_INLET)
        s.b_qy[b] = (m.vol[L]) / (s.ap[L]) * Gas::dot3(m.balpha + 3 * b, m.bS + 3 * b);
    if (t == OCSL_BC_PRESSURE_OUTLET) {
        if (s.b_super[b]) {
            s.b_mdot[b] = g.density(s.p[L], s.T[L]) * Gas::dot3(s.vec3(s.u, L).v, m.bS + 3 * b);
            s.b_qy[b] = 1.0;
        } else {
            const double qy = (m.vol[L]) / (s.ap[L]) * Gas::dot3(m.balpha + 3 * b, m.bS + 3 * b);
            s.b_qy[b] = qy;
            double dvec[3];
            for (int k = 0; k < 3; ++k) dvec[k] = m.bdist[b] * m.be[3 * b + k];
            const double YF = qy * (s.b_p[b] - s.p[L] - 0.5 * Gas::dot3(s.vec3(s.gp, L).v, dvec));
            double mdot = Gas::dot3(s.vec3(s.u, L).v, m.bS + 3 * b) - YF;
            mdot *= s.b_rho[b];
            s.b_mdot[b] = mdot;
        }
    }
}

// correct_face_mdot (cflow_solver_supply.cpp:96-130): threads [0, F) the
// interior faces, [F, F + B) the boundary faces.
__global__ void __launch_bounds__(kBlock) k_correct_mdot(MeshDev m, StateDev s, FlowPhys ph, int blend) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < m.F) {
        const int f = q, L = m.fL[f], R = m.fR[f];
        const double geo = m.fgeo[f], md = s.f_mdot[f];
        const double Tf = conv_interp_d(blend, geo, md, s.T[L], s.T[R]);
        const double pcf = conv_interp_d(blend, geo, md, s.pc[L], s.pc[R]);
        double v = md;
        v += md / s.f_rho[f] * ph.gas.dp_drho(Tf) * pcf;
        v += s.f_qy[f] * s.f_rho[f] * (s.pc[L] - s.pc[R]);
        s.f_mdot[f] = v;
    } else if (q < m.F + m.B) {
        const int b = q - m.F;
        if (ph.bc[m.bmark[b]].type == OCSL_BC_PRESSURE_OUTLET && !s.b_super[b])
            s.b_mdot[b] += s.b_qy[b] * s.b_rho[b] * (s.pc[m.bL[b]]);
    }
}

// u_d = the momentum solve's solution (cflow_solver.cpp:290-299).
// x is in the caller's cell order; cperm (null: the same order) maps each
// device cell to it.
__global__ void __launch_bounds__(kBlock) k_take_u(StateDev s, const double* __restrict__ x, const int* cperm, int d,
                                                   int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) s.u[d * s.cs + i] = x[cperm ? cperm[i] : i];
}

// p' = 0.5 x, p += p' (cflow_solver.cpp:350-357).
__global__ void __launch_bounds__(kBlock) k_take_pc(StateDev s, const double* __restrict__ x, const int* cperm,
                                                    int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pc = kRelaxP * x[cperm ? cperm[i] : i];
    s.pc[i] = pc;
    s.p[i] += pc;
}

// Boundary pressure after the pressure update (cflow_solver.cpp:361-374).
__global__ void __launch_bounds__(kBlock) k_bp_after_p(MeshDev m, StateDev s, FlowPhys ph) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.B) return;
    const int t = ph.bc[m.bmark[b]].type;
    if (t == OCSL_BC_VELOCITY_INLET) s.b_p[b] = s.p[m.bL[b]];
    if (t == OCSL_BC_PRESSURE_OUTLET && s.b_super[b]) s.b_p[b] = s.p[m.bL[b]];
}

// Density from the new pressure and the velocity correction
// (cflow_solver.cpp:390-396, 418-426): rho = rho(p, T); u -= 0.8 V grad p' / ap_vec.
__global__ void __launch_bounds__(kBlock) k_rho_u(MeshDev m, StateDev s, FlowPhys ph) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    s.rho[i] = ph.gas.density(s.p[i], s.T[i]);
    const double a = kRelaxU * (m.vol[i]);
    for (int k = 0; k < 3; ++k) {
        double den = s.ap_vec[k * s.cs + i];
        if (den == 0.0) den = 2.2250738585072014e-308;  // DBL_MIN
        s.u[k * s.cs + i] -= a * (s.gpc[k * s.cs + i] / den);
    }
}

// Boundary density after the pressure update (cflow_solver.cpp:401-416).
__global__ void __launch_bounds__(kBlock) k_brho(MeshDev m, StateDev s, FlowPhys ph) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.B) return;
    const Gas& g = ph.gas;
    const int L = m.bL[b], t = ph.bc[m.bmark[b]].type;
    if (t == OCSL_BC_VELOCITY_INLET) s.b_rho[b] = g.density(s.p[L], s.b_T[b]);
    if (t == OCSL_BC_PRESSURE_OUTLET)
        s.b_rho[b] = s.b_super[b] ? g.density(s.p[L], s.T[L]) : g.density(s.b_p[b], s.T[L]);
}

// h = the enthalpy solve's solution, T = T(h, u), Mach (cflow_solver.cpp:462-472).
__global__ void __launch_bounds__(kBlock) k_take_h(StateDev s, Gas g, const double* __restrict__ x, const int* cperm,
                                                   int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double h = x[cperm ? cperm[i] : i];
    s.h[i] = h;
    const V3 ui = s.vec3(s.u, i);
    const double T = g.temperature(h, ui.v);
    s.T[i] = T;
    s.mach[i] = Gas::mag3(ui.v) / g.sound(T);
}

}  // namespace ocsl

using namespace ocsl;

// Host mirror of the flow state (CFlow_cell / CFlow_face / CFlow_BC_face).
struct HostFlow {
    std::vector<double> u, u_old, rho, rho_old, p, p_old, pc, gp, gpc, T, h, h_old, ap, ap_vec, dt, mach;
    std::vector<double> f_mdot, f_rho, f_qy, f_stress;
    std::vector<double> b_T, b_p, b_p_out, b_rho, b_mdot, b_h, b_qy, b_u;
    std::vector<int> b_super;
};

struct HostMesh {
    int n = 0, F = 0, B = 0;
    std::vector<double> vol, fS, fe, falpha, farea, fdist, fksi, fgeo, bS, be, balpha, barea, bdist, bksi, bperp;
    std::vector<int> fL, fR, bL, bmark;
};

struct ocsl_flow_s {
    ocsl_solver* solver = nullptr;
    ocsl_flow_params prm{};
    Gas gas{};
    double urf = 0.5;  // implicit_underrelaxation (cflow_solver.h:177)
    HostMesh mesh;
    HostFlow st;
    // The device keeps interior faces sorted by their left cell, so a cell's
    // faces as owner are contiguous: device face k is the caller's face
    // fperm[k].  The host mirror keeps the caller's order.
    std::vector<int> fperm;
    // Cells: device cell k is the caller's cell cperm[k] (ciperm the inverse);
    // cell_perm false when that is the identity.  Per-cell state and mesh
    // arrays are stored in the device order; interior faces follow their
    // (device) left cell; boundary faces keep the caller's order.
    std::vector<int> cperm, ciperm;
    bool cell_perm = false;
    size_t asm_lds = 0;  // LDS bytes of one assembly block
    DBuf<int> d_fperm, d_cperm, d_rstart, d_rlen;
    DBuf<double> d_fstage;  // caller-order staging of per-face state
    DBuf<double> d_cstage;  // caller-order staging of per-cell state
    bool dev = true;  // support loops on the device, state resident there
    // device mesh
    DBuf<double> d_vol, d_fS, d_fe, d_fdist, d_fksi, d_fgeo, d_falpha, d_bS, d_barea, d_bksi, d_bperp, d_be, d_balpha,
        d_bdist;
    DBuf<int> d_fL, d_fR, d_bL, d_bmark, d_cf_sp, d_cf_code, d_cb_ptr, d_cb_face, d_row_ptr, d_dpos;
    DBuf<int> d_cf_nb;
    // device state
    DBuf<double> d_u, d_u_old, d_rho, d_rho_old, d_p, d_p_old, d_pc, d_gp, d_T, d_h, d_h_old, d_ap, d_ap_vec, d_dt;
    DBuf<double> d_f_mdot, d_f_rho, d_f_qy, d_f_stress;
    DBuf<double> d_b_T, d_b_rho, d_b_mdot, d_b_h, d_b_qy, d_b_u;
    DBuf<int> d_b_super;
    DBuf<double> d_gpc, d_mach, d_b_p, d_b_p_out;
    // assembled system
    DBuf<double> d_val, d_b, d_x0;
    DBuf<double> d_b3, d_x3;  // momentum right-hand sides and solutions, solved together
    int64_t nnz_in = 0;
    std::vector<double> x_host, hist;
    bool initialized = false;

    MeshDev mesh_view() const {
        return {mesh.n,     mesh.F,       mesh.B,      d_vol.p,    d_fL.p,      d_fR.p,       d_fS.p,
                d_fe.p,     d_fdist.p,    d_fksi.p,    d_fgeo.p,   d_falpha.p,  d_bL.p,       d_bmark.p,
                d_bS.p,     d_barea.p,    d_bksi.p,    d_bperp.p,  d_be.p,      d_balpha.p,   d_bdist.p,
                d_cf_sp.p,  d_cf_code.p,  d_cb_ptr.p, d_cb_face.p, d_row_ptr.p,  d_dpos.p,
                d_cf_nb.p,  cell_perm ? d_cperm.p : nullptr, d_rstart.p, d_rlen.p};
    }
    StateDev state_view() const {
        return {d_u.p, d_u_old.p, d_rho.p, d_rho_old.p, d_p.p, d_p_old.p, d_pc.p, d_gp.p, d_T.p, d_h.p,
                d_h_old.p, d_ap.p, d_ap_vec.p, d_dt.p, d_f_mdot.p, d_f_rho.p, d_f_qy.p, d_f_stress.p,
                d_b_T.p, d_b_rho.p, d_b_mdot.p, d_b_h.p, d_b_qy.p, d_b_u.p, d_b_super.p,
                d_gpc.p, d_mach.p, d_b_p.p, d_b_p_out.p, mesh.n, mesh.F};
    }
    FlowPhys phys() const {
        FlowPhys ph{};
        ph.gas = gas;
        ph.urf = urf;
        ph.n_bc = prm.n_bc;
        for (int k = 0; k < prm.n_bc; ++k) ph.bc[k] = {prm.bc[k].type, prm.bc[k].no_slip, 0, 0};
        return ph;
    }
    int bc_type(int bf) const { return prm.bc[mesh.bmark[bf]].type; }
};

namespace {

template <class F>
ocsl_status guarded_f(F&& f) {
    try {
        f();
        return OCSL_OK;
    } catch (const Error& e) {
        set_last_error(e.what);
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return OCSL_ERR_NOMEM;
    }
}

template <class T>
void up(DBuf<T>& d, const std::vector<T>& h, hipStream_t s) {
    d.upload(h.data(), h.size(), s);
}

template <class T>
void down(std::vector<T>& h, const DBuf<T>& d, hipStream_t s) {
    if (!h.empty()) OCSL_HIP(hipMemcpyAsync(h.data(), d.p, sizeof(T) * h.size(), hipMemcpyDeviceToHost, s));
}

// Per-face or per-cell data between the caller's order and the handle's:
// device item k is the caller's item perm[k] (perm null: item k); c
// components per item, interleaved on the caller's side, in planes of F
// items on the device.
template <bool kToHandle>
__global__ void k_perm_items(int F, int c, const int* __restrict__ perm, const double* __restrict__ in,
                             double* __restrict__ out) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < F; k += gridDim.x * blockDim.x) {
        const size_t f = perm ? (size_t)perm[k] : (size_t)k;
        for (int q = 0; q < c; ++q) {
            if (kToHandle) out[(size_t)q * F + k] = in[f * c + q];
            else out[f * c + q] = in[(size_t)q * F + k];
        }
    }
}

// Set-up upload of a per-face mesh array, permuted on the host.
template <class T>
void up_mesh_f(const ocsl_flow_s* fl, DBuf<T>& d, const std::vector<T>& h, hipStream_t s) {
    const size_t F = fl->fperm.size(), c = F ? h.size() / F : 0;
    std::vector<T> t(h.size());
    for (size_t k = 0; k < F; ++k)
        for (size_t q = 0; q < c; ++q) t[k * c + q] = h[(size_t)fl->fperm[k] * c + q];
    d.upload(t.data(), t.size(), s);
    OCSL_HIP(hipStreamSynchronize(s));  // t is freed on return
}

// Per-face state (1 or 3 components per face) from the host mirror, in the
// caller's order, to the device, in the handle's order: staged, then gathered.
void up_f(ocsl_flow_s* fl, DBuf<double>& d, const std::vector<double>& h, hipStream_t s) {
    const int F = (int)fl->fperm.size();
    d.alloc(h.size());
    if (h.empty() || F == 0) return;
    fl->d_fstage.upload(h.data(), h.size(), s);
    k_perm_items<true><<<grid_for(F), kBlock, 0, s>>>(F, (int)(h.size() / F), fl->d_fperm.p, fl->d_fstage.p, d.p);
    OCSL_HIP(hipGetLastError());
}

// The reverse: scattered into the staging buffer, then copied out (ordered
// on the stream; the caller synchronises before reading h).
void down_f(ocsl_flow_s* fl, std::vector<double>& h, const DBuf<double>& d, hipStream_t s) {
    const int F = (int)fl->fperm.size();
    if (h.empty() || F == 0) return;
    fl->d_fstage.alloc(h.size());
    k_perm_items<false><<<grid_for(F), kBlock, 0, s>>>(F, (int)(h.size() / F), fl->d_fperm.p, d.p, fl->d_fstage.p);
    OCSL_HIP(hipGetLastError());
    OCSL_HIP(hipMemcpyAsync(h.data(), fl->d_fstage.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
}

// Per-cell state (1 or 3 components per cell) between the host mirror, in the
// caller's order with interleaved vectors, and the device, in the handle's
// order with vectors in planes.
void up_c(ocsl_flow_s* fl, DBuf<double>& d, const std::vector<double>& h, hipStream_t s) {
    const int n = fl->mesh.n;
    if (!fl->cell_perm && h.size() <= (size_t)n) {
        up(d, h, s);
        return;
    }
    d.alloc(h.size());
    if (h.empty()) return;
    fl->d_cstage.upload(h.data(), h.size(), s);
    k_perm_items<true><<<grid_for(n), kBlock, 0, s>>>(n, (int)(h.size() / n), fl->cell_perm ? fl->d_cperm.p : nullptr,
                                                       fl->d_cstage.p, d.p);
    OCSL_HIP(hipGetLastError());
}

void down_c(ocsl_flow_s* fl, std::vector<double>& h, const DBuf<double>& d, hipStream_t s) {
    const int n = fl->mesh.n;
    if (!fl->cell_perm && h.size() <= (size_t)n) {
        down(h, d, s);
        return;
    }
    if (h.empty()) return;
    fl->d_cstage.alloc(h.size());
    k_perm_items<false><<<grid_for(n), kBlock, 0, s>>>(n, (int)(h.size() / n), fl->cell_perm ? fl->d_cperm.p : nullptr,
                                                        d.p, fl->d_cstage.p);
    OCSL_HIP(hipGetLastError());
    OCSL_HIP(hipMemcpyAsync(h.data(), fl->d_cstage.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
}

inline const double* v3(const std::vector<double>& a, int i) { return a.data() + 3 * i; }

// ---- host restatement of the SIMPLE-step support (cflow_solver_supply.cpp) --

template <bool kBlend>
double conv_interp_h(double geo, double mdot, double l, double r) {
    auto avg = [&](double x, double y) {
        double v = geo * x;
        v += (1.0 - geo) * y;
        return v;
    };
    if (mdot >= 0.0) return kBlend ? l * kUpwind + kAlpha * avg(l, r) : l;
    return kBlend ? r * kUpwind + kAlpha * avg(l, r) : r;
}

double conv_interp_rt(bool blend, double geo, double mdot, double l, double r) {
    return blend ? conv_interp_h<true>(geo, mdot, l, r) : conv_interp_h<false>(geo, mdot, l, r);
}

// calc_face_mdot, interior faces (cflow_solver_supply.cpp:25-46)
void face_mdot(ocsl_flow_s* fl) {
    const HostMesh& m = fl->mesh;
    HostFlow& s = fl->st;
    for (int f = 0; f < m.F; ++f) {
        const int L = m.fL[f], R = m.fR[f];
        const double geo = m.fgeo[f];
        s.f_qy[f] = ((m.vol[L] / s.ap[L]) + (m.vol[R] / s.ap[R])) * Gas::dot3(v3(m.falpha, f), v3(m.fS, f));
        double pgf[3], vf[3], dvec[3];
        for (int k = 0; k < 3; ++k) {
            pgf[k] = geo * s.gp[3 * L + k];
            pgf[k] += (1.0 - geo) * s.gp[3 * R + k];
            vf[k] = geo * s.u[3 * L + k];
            vf[k] += (1.0 - geo) * s.u[3 * R + k];
            dvec[k] = m.fdist[f] * m.fe[3 * f + k];
        }
        const double YF = s.f_qy[f] * (s.p[R] - s.p[L] - Gas::dot3(pgf, dvec));
        double mdot = Gas::dot3(vf, v3(m.fS, f)) - YF;
        s.f_rho[f] = conv_interp_rt(fl->prm.blended, geo, mdot, s.rho[L], s.rho[R]);
        mdot *= s.f_rho[f];
        s.f_mdot[f] = mdot;
    }
}

// calc_face_mdot, boundary faces (cflow_solver_supply.cpp:54-93)
void bc_mdot(ocsl_flow_s* fl) {
    const HostMesh& m = fl->mesh;
    HostFlow& s = fl->st;
    const Gas& g = fl->gas;
    for (int b = 0; b < m.B; ++b) {
        const int L = m.bL[b], t = fl->bc_type(b);
        if (t == OCSL_BC_VELOCITY_INLET)
            s.b_mdot[b] = g.density(s.p[L], s.b_T[b]) * Gas::dot3(v3(s.b_u, b), v3(m.bS, b));
        if (t == OCSL_BC_SUPERSONIC_INLET)
            s.b_qy[b] = (m.vol[L]) / (s.ap[L]) * Gas::dot3(v3(m.balpha, b), v3(m.bS, b));
        if (t == OCSL_BC_PRESSURE_OUTLET) {
            if (s.b_super[b]) {
                s.b_mdot[b] = g.density(s.p[L], s.T[L]) * Gas::dot3(v3(s.u, L), v3(m.bS, b));
                s.b_qy[b] = 1.0;
            } else {
                s.b_qy[b] = (m.vol[L]) / (s.ap[L]) * Gas::dot3(v3(m.balpha, b), v3(m.bS, b));
                double dvec[3];
                for (int k = 0; k < 3; ++k) dvec[k] = m.bdist[b] * m.be[3 * b + k];
                const double YF = s.b_qy[b] * (s.b_p[b] - s.p[L] - 0.5 * Gas::dot3(v3(s.gp, L), dvec));
                double mdot = Gas::dot3(v3(s.u, L), v3(m.bS, b)) - YF;
                mdot *= s.b_rho[b];
                s.b_mdot[b] = mdot;
            }
        }
    }
}

// correct_face_mdot (cflow_solver_supply.cpp:96-130)
void correct_mdot(ocsl_flow_s* fl) {
    const HostMesh& m = fl->mesh;
    HostFlow& s = fl->st;
    const Gas& g = fl->gas;
    const bool blend = fl->prm.blended;
    for (int f = 0; f < m.F; ++f) {
        const int L = m.fL[f], R = m.fR[f];
        const double geo = m.fgeo[f];
        const double Tf = conv_interp_rt(blend, geo, s.f_mdot[f], s.T[L], s.T[R]);
        const double pcf = conv_interp_rt(blend, geo, s.f_mdot[f], s.pc[L], s.pc[R]);
        s.f_mdot[f] += s.f_mdot[f] / s.f_rho[f] * g.dp_drho(Tf) * pcf;
        s.f_mdot[f] += s.f_qy[f] * s.f_rho[f] * (s.pc[L] - s.pc[R]);
    }
    for (int b = 0; b < m.B; ++b)
        if (fl->bc_type(b) == OCSL_BC_PRESSURE_OUTLET && !s.b_super[b])
            s.b_mdot[b] += s.b_qy[b] * s.b_rho[b] * (s.pc[m.bL[b]]);
}

// calc_press_grad (cflow_solver_supply.cpp:132-175)
void press_grad(ocsl_flow_s* fl) {
    const HostMesh& m = fl->mesh;
    HostFlow& s = fl->st;
    std::fill(s.gp.begin(), s.gp.end(), 0.0);
    for (int f = 0; f < m.F; ++f) {
        const int L = m.fL[f], R = m.fR[f];
        double fp = (s.ap[L] + s.ap[R]);
        fp = (s.ap[L] * s.p[L] + s.ap[R] * s.p[R]) / fp;
        for (int k = 0; k < 3; ++k) {
            s.gp[3 * L + k] = s.gp[3 * L + k] + fp * m.fS[3 * f + k];
            s.gp[3 * R + k] = s.gp[3 * R + k] - fp * m.fS[3 * f + k];
        }
    }
    for (int b = 0; b < m.B; ++b) {
        const int L = m.bL[b], t = fl->bc_type(b);
        double pv;
        if (t == OCSL_BC_WALL || t == OCSL_BC_SYMMETRY || t == OCSL_BC_EMPTY || t == OCSL_BC_VELOCITY_INLET)
            pv = s.p[L];
        else if (t == OCSL_BC_PRESSURE_OUTLET || t == OCSL_BC_SUPERSONIC_INLET)
            pv = s.b_p[b];
        else
            continue;
        for (int k = 0; k < 3; ++k) s.gp[3 * L + k] = s.gp[3 * L + k] + pv * m.bS[3 * b + k];
    }
    for (int i = 0; i < m.n; ++i)
        for (int k = 0; k < 3; ++k) s.gp[3 * i + k] = (1.0 / m.vol[i]) * s.gp[3 * i + k];
}

// calc_press_corr_grad (cflow_solver_supply.cpp:177-214)
void press_corr_grad(ocsl_flow_s* fl) {
    const HostMesh& m = fl->mesh;
    HostFlow& s = fl->st;
    std::fill(s.gpc.begin(), s.gpc.end(), 0.0);
    for (int f = 0; f < m.F; ++f) {
        const int L = m.fL[f], R = m.fR[f];
        double fp = m.fgeo[f] * s.pc[L];
        fp += (1.0 - m.fgeo[f]) * s.pc[R];
        for (int k = 0; k < 3; ++k) {
            s.gpc[3 * L + k] = s.gpc[3 * L + k] + fp * m.fS[3 * f + k];
            s.gpc[3 * R + k] = s.gpc[3 * R + k] - fp * m.fS[3 * f + k];
        }
    }
    for (int b = 0; b < m.B; ++b) {
        const int L = m.bL[b], t = fl->bc_type(b);
        if (t == OCSL_BC_WALL || t == OCSL_BC_SYMMETRY || t == OCSL_BC_EMPTY || t == OCSL_BC_VELOCITY_INLET ||
            (s.b_super[b] && t == OCSL_BC_PRESSURE_OUTLET))
            for (int k = 0; k < 3; ++k) s.gpc[3 * L + k] = s.gpc[3 * L + k] + s.pc[L] * m.bS[3 * b + k];
    }
    for (int i = 0; i < m.n; ++i)
        for (int k = 0; k < 3; ++k) s.gpc[3 * i + k] = (1.0 / m.vol[i]) * s.gpc[3 * i + k];
}

// set_boundary_conditions (cflow_solver.cpp:192-227)
void set_bcs(ocsl_flow_s* fl) {
    const HostMesh& m = fl->mesh;
    HostFlow& s = fl->st;
    const Gas& g = fl->gas;
    for (int b = 0; b < m.B; ++b) {
        const int L = m.bL[b], t = fl->bc_type(b);
        if (t == OCSL_BC_VELOCITY_INLET) {
            s.b_p[b] = s.p[L];
            s.b_rho[b] = g.density(s.p[L], s.b_T[b]);
            s.b_mdot[b] = s.b_rho[b] * Gas::dot3(v3(s.b_u, b), v3(m.bS, b));
            s.b_h[b] = g.enthalpy(s.b_T[b], v3(s.b_u, b));
        }
        if (t == OCSL_BC_PRESSURE_OUTLET) {
            s.b_T[b] = s.T[L];
            for (int k = 0; k < 3; ++k) s.b_u[3 * b + k] = s.u[3 * L + k];
            s.b_rho[b] = g.density(s.b_p[b], s.T[L]);
            s.b_h[b] = g.enthalpy(s.T[L], v3(s.u, L));
            // both branches mark the face supersonic (cflow_solver.cpp:208-217)
            if ((s.b_mdot[b] / s.rho[L] / m.barea[b]) >= g.sound(s.T[L])) {
                s.b_super[b] = 1;
                s.b_p[b] = s.p[L];
            } else {
                s.b_super[b] = 1;
                s.b_p[b] = s.b_p_out[b];
            }
        }
        if (t == OCSL_BC_WALL)
            s.b_p[b] = s.p[L] + Gas::dot3(v3(s.gp, L), v3(m.be, b)) * m.bdist[b];
    }
}

void upload_cells(ocsl_flow_s* fl, int eq) {
    HostFlow& s = fl->st;
    hipStream_t st = fl->solver->stream;
    up_c(fl, fl->d_T, s.T, st);
    up_c(fl, fl->d_rho, s.rho, st);
    up_c(fl, fl->d_dt, s.dt, st);
    up_f(fl, fl->d_f_mdot, s.f_mdot, st);
    up(fl->d_b_mdot, s.b_mdot, st);
    if (eq <= OCSL_EQ_MOMENTUM_Z) {
        up_c(fl, fl->d_u, s.u, st);
        up_c(fl, fl->d_u_old, s.u_old, st);
        up_c(fl, fl->d_rho_old, s.rho_old, st);
        up_c(fl, fl->d_gp, s.gp, st);
        up_c(fl, fl->d_ap, s.ap, st);
        up_c(fl, fl->d_ap_vec, s.ap_vec, st);
        up_f(fl, fl->d_f_stress, s.f_stress, st);
        up(fl->d_b_u, s.b_u, st);
    } else if (eq == OCSL_EQ_PRESSURE) {
        up_c(fl, fl->d_rho_old, s.rho_old, st);
        up_c(fl, fl->d_pc, s.pc, st);
        up_f(fl, fl->d_f_rho, s.f_rho, st);
        up_f(fl, fl->d_f_qy, s.f_qy, st);
        up(fl->d_b_T, s.b_T, st);
        up(fl->d_b_rho, s.b_rho, st);
        up(fl->d_b_qy, s.b_qy, st);
        up(fl->d_b_super, s.b_super, st);
    } else {
        up_c(fl, fl->d_h, s.h, st);
        up_c(fl, fl->d_h_old, s.h_old, st);
        up_c(fl, fl->d_rho_old, s.rho_old, st);
        up_c(fl, fl->d_p, s.p, st);
        up_c(fl, fl->d_p_old, s.p_old, st);
        up(fl->d_b_h, s.b_h, st);
        up(fl->d_b_T, s.b_T, st);
    }
}

// The whole state between host mirror and device (device-resident mode).
void upload_all(ocsl_flow_s* fl) {
    const HostFlow& s = fl->st;
    hipStream_t st = fl->solver->stream;
    up_c(fl, fl->d_u, s.u, st);
    up_c(fl, fl->d_u_old, s.u_old, st);
    up_c(fl, fl->d_rho, s.rho, st);
    up_c(fl, fl->d_rho_old, s.rho_old, st);
    up_c(fl, fl->d_p, s.p, st);
    up_c(fl, fl->d_p_old, s.p_old, st);
    up_c(fl, fl->d_pc, s.pc, st);
    up_c(fl, fl->d_gp, s.gp, st);
    up_c(fl, fl->d_gpc, s.gpc, st);
    up_c(fl, fl->d_T, s.T, st);
    up_c(fl, fl->d_h, s.h, st);
    up_c(fl, fl->d_h_old, s.h_old, st);
    up_c(fl, fl->d_ap, s.ap, st);
    up_c(fl, fl->d_ap_vec, s.ap_vec, st);
    up_c(fl, fl->d_dt, s.dt, st);
    up_c(fl, fl->d_mach, s.mach, st);
    up_f(fl, fl->d_f_mdot, s.f_mdot, st);
    up_f(fl, fl->d_f_rho, s.f_rho, st);
    up_f(fl, fl->d_f_qy, s.f_qy, st);
    up_f(fl, fl->d_f_stress, s.f_stress, st);
    up(fl->d_b_T, s.b_T, st);
    up(fl->d_b_p, s.b_p, st);
    up(fl->d_b_p_out, s.b_p_out, st);
    up(fl->d_b_rho, s.b_rho, st);
    up(fl->d_b_mdot, s.b_mdot, st);
    up(fl->d_b_h, s.b_h, st);
    up(fl->d_b_qy, s.b_qy, st);
    up(fl->d_b_u, s.b_u, st);
    up(fl->d_b_super, s.b_super, st);
    OCSL_HIP(hipStreamSynchronize(st));
}

void download_all(ocsl_flow_s* fl) {
    HostFlow& s = fl->st;
    hipStream_t st = fl->solver->stream;
    down_c(fl, s.u, fl->d_u, st);
    down_c(fl, s.u_old, fl->d_u_old, st);
    down_c(fl, s.rho, fl->d_rho, st);
    down_c(fl, s.rho_old, fl->d_rho_old, st);
    down_c(fl, s.p, fl->d_p, st);
    down_c(fl, s.p_old, fl->d_p_old, st);
    down_c(fl, s.pc, fl->d_pc, st);
    down_c(fl, s.gp, fl->d_gp, st);
    down_c(fl, s.gpc, fl->d_gpc, st);
    down_c(fl, s.T, fl->d_T, st);
    down_c(fl, s.h, fl->d_h, st);
    down_c(fl, s.h_old, fl->d_h_old, st);
    down_c(fl, s.ap, fl->d_ap, st);
    down_c(fl, s.ap_vec, fl->d_ap_vec, st);
    down_c(fl, s.dt, fl->d_dt, st);
    down_c(fl, s.mach, fl->d_mach, st);
    down_f(fl, s.f_mdot, fl->d_f_mdot, st);
    down_f(fl, s.f_rho, fl->d_f_rho, st);
    down_f(fl, s.f_qy, fl->d_f_qy, st);
    down_f(fl, s.f_stress, fl->d_f_stress, st);
    down(s.b_T, fl->d_b_T, st);
    down(s.b_p, fl->d_b_p, st);
    down(s.b_p_out, fl->d_b_p_out, st);
    down(s.b_rho, fl->d_b_rho, st);
    down(s.b_mdot, fl->d_b_mdot, st);
    down(s.b_h, fl->d_b_h, st);
    down(s.b_qy, fl->d_b_qy, st);
    down(s.b_u, fl->d_b_u, st);
    down(s.b_super, fl->d_b_super, st);
    OCSL_HIP(hipStreamSynchronize(st));
}

template <int kEq>
void launch_assemble(ocsl_flow_s* fl, int d) {
    const ocsl_flow_params& p = fl->prm;
    const int n = fl->mesh.n, grid = (n + kBlock - 1) / kBlock;
    hipStream_t st = fl->solver->stream;
    const MeshDev m = fl->mesh_view();
    const StateDev s = fl->state_view();
    const FlowPhys ph = fl->phys();
    double *val = fl->d_val.p, *b = fl->d_b.p, *x0 = fl->d_x0.p;
#define OCSL_ASM(V, BL, TR)                                                                       \
    do {                                                                                          \
        if (fl->cell_perm)                                                                        \
            k_assemble<kEq, V, BL, TR, true><<<grid, kBlock, fl->asm_lds, st>>>(m, s, ph, d, val, b, x0);  \
        else                                                                                      \
            k_assemble<kEq, V, BL, TR, false><<<grid, kBlock, fl->asm_lds, st>>>(m, s, ph, d, val, b, x0); \
    } while (0)
    const int key = (p.viscous ? 4 : 0) + (p.blended ? 2 : 0) + (p.transient ? 1 : 0);
    switch (key) {
        case 0: OCSL_ASM(false, false, false); break;
        case 1: OCSL_ASM(false, false, true); break;
        case 2: OCSL_ASM(false, true, false); break;
        case 3: OCSL_ASM(false, true, true); break;
        case 4: OCSL_ASM(true, false, false); break;
        case 5: OCSL_ASM(true, false, true); break;
        case 6: OCSL_ASM(true, true, false); break;
        default: OCSL_ASM(true, true, true); break;
    }
#undef OCSL_ASM
    OCSL_HIP(hipGetLastError());
}

// Assembles equation eq into d_val / d_b / d_x0 and hands them to the solver.
void assemble(ocsl_flow_s* fl, int eq) {
    OCSL_REQUIRE(eq >= OCSL_EQ_MOMENTUM_X && eq <= OCSL_EQ_ENTHALPY, OCSL_ERR_ARG, "bad equation");
    hipStream_t st = fl->solver->stream;
    if (!fl->dev) upload_cells(fl, eq);
    if (eq <= OCSL_EQ_MOMENTUM_Z) launch_assemble<kMomentum>(fl, eq);
    else if (eq == OCSL_EQ_PRESSURE) launch_assemble<kPressure>(fl, 0);
    else launch_assemble<kEnthalpy>(fl, 0);
    if (eq <= OCSL_EQ_MOMENTUM_Z && !fl->dev) {
        down_c(fl, fl->st.ap, fl->d_ap, st);
        down_c(fl, fl->st.ap_vec, fl->d_ap_vec, st);
    }
    if (!fl->dev) OCSL_HIP(hipStreamSynchronize(st));  // the host reads ap, ap_vec
    ocsl_solver* h = fl->solver;
    h->val_in = fl->d_val.p;
    h->have_vals = true;
    h->val_epoch++;
    h->csr_val_valid = h->csr_val32_valid = false;
    h->b_dev = fl->d_b.p;
    h->have_b = true;
    h->x_own.alloc(fl->mesh.n);
    OCSL_HIP(hipMemcpyAsync(h->x_own.p, fl->d_x0.p, sizeof(double) * fl->mesh.n, hipMemcpyDeviceToDevice, st));
}

void viscous_stress(ocsl_flow_s* fl) {
    hipStream_t st = fl->solver->stream;
    if (!fl->dev) {
        up_c(fl, fl->d_u, fl->st.u, st);
        up_c(fl, fl->d_T, fl->st.T, st);
    }
    const int F = fl->mesh.F;
    if (F > 0)
        k_visc_stress<<<(F + kBlock - 1) / kBlock, kBlock, 0, st>>>(fl->mesh_view(), fl->state_view(), fl->gas);
    OCSL_HIP(hipGetLastError());
    if (!fl->dev) {
        down_f(fl, fl->st.f_stress, fl->d_f_stress, st);
        OCSL_HIP(hipStreamSynchronize(st));
    }
}

// Solves the assembled system with the configured method; the solution lands
// in fl->x_host.
void solve(ocsl_flow_s* fl, int method) {
    ocsl_solver* h = fl->solver;
    double res = 0.0;
    if (fl->prm.mode == OCSL_MODE_REFERENCE) {
        const ocsl_status st = method == OCSL_CG ? ocsl_CG_solve(h, &res) : ocsl_ILU_BICGSTAB_solve(h, &res);
        if (st < 0) throw Error{st, std::string("flow solve: ") + ocsl_last_error()};
    } else {
        ocsl_report rep{};
        const ocsl_status st = ocsl_solve(h, method, OCSL_MODE_TOLERANCE, fl->prm.precision, &fl->prm.solve_options,
                                          &rep);
        if (st < 0) throw Error{st, std::string("flow solve: ") + ocsl_last_error()};
        res = rep.residual;
    }
    fl->hist.push_back(res);
    if (fl->dev) return;  // the solution stays in the handle's x for the device kernels
    fl->x_host.resize(fl->mesh.n);
    const ocsl_status st = ocsl_get_x(h, fl->x_host.data(), OCSL_HOST);
    if (st < 0) throw Error{st, std::string("flow solve: ") + ocsl_last_error()};
}

// The three momentum systems share one matrix and none reads another's
// solution unless a viscous flow has a no-slip wall, whose terms couple the
// components (cflow_solver_cell_face_loops.cpp:94-113); then one REFERENCE
// solve of three right-hand sides replaces the three consecutive ones.
bool momentum_together(const ocsl_flow_s* fl) {
    const ocsl_flow_params& p = fl->prm;
    if (!p.momentum_together || p.mode != OCSL_MODE_REFERENCE || p.method_momentum != OCSL_ILU_BICGSTAB) return false;
    if (!p.viscous) return true;
    for (int k = 0; k < p.n_bc; ++k)
        if (p.bc[k].type == OCSL_BC_WALL && p.bc[k].no_slip) return false;
    return true;
}

// CFlow_solver::iterate (cflow_solver.cpp:229-476) with the state resident in
// device memory: the support loops as kernels on the solver's stream, the
// solutions read from the handle's x.
void iterate_dev(ocsl_flow_s* fl) {
    const int n = fl->mesh.n, F = fl->mesh.F, B = fl->mesh.B;
    hipStream_t st = fl->solver->stream;
    const MeshDev m = fl->mesh_view();
    const StateDev s = fl->state_view();
    const FlowPhys ph = fl->phys();
    const int blend = fl->prm.blended ? 1 : 0;
    auto grid = [](int cnt) { return (cnt + kBlock - 1) / kBlock; };
    const double* x = nullptr;
    if (B) k_set_bcs<<<grid(B), kBlock, 0, st>>>(m, s, ph);
    k_press_grad<<<grid(n), kBlock, 0, st>>>(m, s, ph);
    k_fill<double><<<grid_for(n), kBlock, 0, st>>>(n, 0.0, s.ap);
    OCSL_HIP(hipGetLastError());
    if (fl->prm.viscous) viscous_stress(fl);
    if (momentum_together(fl)) {
        // each component's right-hand side and guess into a column of B3, X3
        double* B3 = (fl->d_b3.alloc(3 * (size_t)n), fl->d_b3.p);
        double* X3 = (fl->d_x3.alloc(3 * (size_t)n), fl->d_x3.p);
        const size_t bytes = sizeof(double) * n;
        for (int d = 0; d < 3; ++d) {
            assemble(fl, OCSL_EQ_MOMENTUM_X + d);
            OCSL_HIP(hipMemcpyAsync(B3 + (size_t)d * n, fl->d_b.p, bytes, hipMemcpyDeviceToDevice, st));
            OCSL_HIP(hipMemcpyAsync(X3 + (size_t)d * n, fl->d_x0.p, bytes, hipMemcpyDeviceToDevice, st));
        }
        ocsl_report reps[3];
        bicgstab_reference_multi(fl->solver, 3, B3, X3, reps);
        for (int d = 0; d < 3; ++d) {
            fl->hist.push_back(reps[d].residual);
            k_take_u<<<grid(n), kBlock, 0, st>>>(s, X3 + (size_t)d * n, m.cperm, d, n);
        }
    } else {
        for (int d = 0; d < 3; ++d) {
            assemble(fl, OCSL_EQ_MOMENTUM_X + d);
            solve(fl, fl->prm.method_momentum);
            x = fl->solver->x_own.p;
            k_take_u<<<grid(n), kBlock, 0, st>>>(s, x, m.cperm, d, n);
        }
    }
    if (F) k_face_mdot<<<grid(F), kBlock, 0, st>>>(m, s, blend);
    if (B) k_bc_mdot<<<grid(B), kBlock, 0, st>>>(m, s, ph);
    OCSL_HIP(hipGetLastError());
    assemble(fl, OCSL_EQ_PRESSURE);
    solve(fl, fl->prm.method_pressure);
    x = fl->solver->x_own.p;
    k_take_pc<<<grid(n), kBlock, 0, st>>>(s, x, m.cperm, n);
    if (B) k_bp_after_p<<<grid(B), kBlock, 0, st>>>(m, s, ph);
    k_press_corr_grad<<<grid(n), kBlock, 0, st>>>(m, s, ph);
    if (F + B) k_correct_mdot<<<grid(F + B), kBlock, 0, st>>>(m, s, ph, blend);
    k_rho_u<<<grid(n), kBlock, 0, st>>>(m, s, ph);
    if (B) k_brho<<<grid(B), kBlock, 0, st>>>(m, s, ph);
    OCSL_HIP(hipGetLastError());
    assemble(fl, OCSL_EQ_ENTHALPY);
    solve(fl, fl->prm.method_enthalpy);
    x = fl->solver->x_own.p;
    k_take_h<<<grid(n), kBlock, 0, st>>>(s, fl->gas, x, m.cperm, n);
    OCSL_HIP(hipGetLastError());
}

// CFlow_solver::iterate (cflow_solver.cpp:229-476)
void iterate(ocsl_flow_s* fl) {
    if (fl->dev) {
        iterate_dev(fl);
        return;
    }
    const HostMesh& m = fl->mesh;
    HostFlow& s = fl->st;
    const Gas& g = fl->gas;
    const int n = m.n;
    set_bcs(fl);
    press_grad(fl);
    std::fill(s.ap.begin(), s.ap.end(), 0.0);
    if (fl->prm.viscous) viscous_stress(fl);
    for (int d = 0; d < 3; ++d) {
        assemble(fl, OCSL_EQ_MOMENTUM_X + d);
        solve(fl, fl->prm.method_momentum);
        for (int i = 0; i < n; ++i) s.u[3 * i + d] = fl->x_host[i];
    }
    face_mdot(fl);
    bc_mdot(fl);
    assemble(fl, OCSL_EQ_PRESSURE);
    solve(fl, fl->prm.method_pressure);
    for (int i = 0; i < n; ++i) {
        s.pc[i] = kRelaxP * fl->x_host[i];
        s.p[i] += s.pc[i];
    }
    for (int b = 0; b < m.B; ++b) {
        const int t = fl->bc_type(b);
        if (t == OCSL_BC_VELOCITY_INLET) s.b_p[b] = s.p[m.bL[b]];
        if (t == OCSL_BC_PRESSURE_OUTLET && s.b_super[b]) s.b_p[b] = s.p[m.bL[b]];
    }
    press_corr_grad(fl);
    correct_mdot(fl);
    for (int i = 0; i < n; ++i) s.rho[i] = g.density(s.p[i], s.T[i]);
    for (int b = 0; b < m.B; ++b) {
        const int L = m.bL[b], t = fl->bc_type(b);
        if (t == OCSL_BC_VELOCITY_INLET) s.b_rho[b] = g.density(s.p[L], s.b_T[b]);
        if (t == OCSL_BC_PRESSURE_OUTLET)
            s.b_rho[b] = s.b_super[b] ? g.density(s.p[L], s.T[L]) : g.density(s.b_p[b], s.T[L]);
    }
    for (int i = 0; i < n; ++i) {
        const double a = kRelaxU * (m.vol[i]);
        for (int k = 0; k < 3; ++k) {
            double den = s.ap_vec[3 * i + k];
            if (den == 0.0) den = std::numeric_limits<double>::min();
            s.u[3 * i + k] -= a * (s.gpc[3 * i + k] / den);
        }
    }
    assemble(fl, OCSL_EQ_ENTHALPY);
    solve(fl, fl->prm.method_enthalpy);
    for (int i = 0; i < n; ++i) {
        s.h[i] = fl->x_host[i];
        s.T[i] = g.temperature(s.h[i], v3(s.u, i));
        s.mach[i] = Gas::mag3(v3(s.u, i)) / g.sound(s.T[i]);
    }
}

// update_timestep (cflow_solver.cpp:497-519)
void update_timestep(ocsl_flow_s* fl) {
    HostFlow& s = fl->st;
    if (fl->dev) {
        hipStream_t st = fl->solver->stream;
        const size_t n = fl->mesh.n;
        const auto d2d = hipMemcpyDeviceToDevice;
        OCSL_HIP(hipMemcpyAsync(fl->d_u_old.p, fl->d_u.p, sizeof(double) * 3 * n, d2d, st));
        OCSL_HIP(hipMemcpyAsync(fl->d_rho_old.p, fl->d_rho.p, sizeof(double) * n, d2d, st));
        OCSL_HIP(hipMemcpyAsync(fl->d_p_old.p, fl->d_p.p, sizeof(double) * n, d2d, st));
        OCSL_HIP(hipMemcpyAsync(fl->d_h_old.p, fl->d_h.p, sizeof(double) * n, d2d, st));
    } else {
        s.u_old = s.u;
        s.rho_old = s.rho;
        s.p_old = s.p;
        s.h_old = s.h;
    }
    const ocsl_flow_params& p = fl->prm;
    if (!p.transient && fl->urf < p.implicit_underrelax_max)
        fl->urf += (p.implicit_underrelax_max - p.implicit_underrelax) / p.ramp_iterations;
}

// CFlow_solver::initialize (cflow_solver.cpp:34-190)
void initialize(ocsl_flow_s* fl) {
    const HostMesh& m = fl->mesh;
    HostFlow& s = fl->st;
    const Gas& g = fl->gas;
    const ocsl_flow_params& p = fl->prm;
    const int n = m.n;
    for (int i = 0; i < n; ++i) {
        s.T[i] = p.t0;
        s.p[i] = p.p0;
        for (int k = 0; k < 3; ++k) s.u[3 * i + k] = p.u0[k];
        s.rho[i] = g.density(s.p[i], s.T[i]);
        s.h[i] = g.enthalpy(s.T[i], v3(s.u, i));
        s.rho_old[i] = s.rho[i];
        s.p_old[i] = s.p[i];
        for (int k = 0; k < 3; ++k) s.u_old[3 * i + k] = s.u[3 * i + k];
        s.h_old[i] = s.h[i];
    }
    if (p.shock_tube) {
        for (int i = n / 2 - 1; i < n; ++i) {
            s.p[i] = p.p2;
            s.T[i] = p.t2;
            s.rho[i] = g.density(s.p[i], s.T[i]);
            s.h[i] = g.enthalpy(s.T[i], v3(s.u, i));
            s.rho_old[i] = s.rho[i];
            s.p_old[i] = s.p[i];
            s.h_old[i] = s.h[i];
        }
    }
    for (int b = 0; b < m.B; ++b) {
        const int L = m.bL[b];
        const ocsl_bc& bc = p.bc[m.bmark[b]];
        const int t = bc.type;
        if (t == OCSL_BC_VELOCITY_INLET || t == OCSL_BC_SUPERSONIC_INLET) {
            for (int k = 0; k < 3; ++k) s.b_u[3 * b + k] = bc.velocity[k];
            s.b_p[b] = t == OCSL_BC_VELOCITY_INLET ? s.p[L] : bc.pressure;
            s.b_T[b] = bc.temperature;
            s.b_rho[b] = g.density(s.b_p[b], s.b_T[b]);
            s.b_h[b] = g.enthalpy(s.b_T[b], v3(s.b_u, b));
            s.b_mdot[b] = s.b_rho[b] * Gas::dot3(v3(s.b_u, b), v3(m.bS, b));
        }
        if (t == OCSL_BC_PRESSURE_OUTLET) {
            s.b_T[b] = s.T[L];
            s.b_p[b] = bc.pressure;
            s.b_p_out[b] = bc.pressure;
            s.b_rho[b] = g.density(s.b_p[b], s.b_T[b]);
            for (int k = 0; k < 3; ++k) s.b_u[3 * b + k] = s.u[3 * L + k];
            s.b_mdot[b] = s.rho[L] * Gas::dot3(v3(s.u, L), v3(m.bS, b));
            s.b_h[b] = g.enthalpy(s.T[L], v3(s.u, L));
            if ((s.b_mdot[b] / s.rho[L] / m.barea[b]) >= g.sound(s.T[L])) {
                s.b_super[b] = 1;
                s.b_p[b] = s.p[L];
                s.b_rho[b] = g.density(s.b_p[b], s.b_T[b]);
                s.b_mdot[b] = s.rho[L] * Gas::dot3(v3(s.u, L), v3(m.bS, b));
            }
        }
        if (t == OCSL_BC_WALL) s.b_p[b] = s.p[L];
    }
    std::fill(s.ap.begin(), s.ap.end(), 1.0);
    face_mdot(fl);
    bc_mdot(fl);
    if (!p.transient) fl->urf = p.implicit_underrelax;
    else std::fill(s.dt.begin(), s.dt.end(), p.dt);
    if (fl->dev) upload_all(fl);
    fl->initialized = true;
}

void alloc_state(ocsl_flow_s* fl) {
    const int n = fl->mesh.n, F = fl->mesh.F, B = fl->mesh.B;
    HostFlow& s = fl->st;
    for (auto* v : {&s.u, &s.u_old, &s.gp, &s.gpc, &s.ap_vec}) v->assign(3 * (size_t)n, 0.0);
    for (auto* v : {&s.rho, &s.rho_old, &s.p, &s.p_old, &s.pc, &s.T, &s.h, &s.h_old, &s.mach}) v->assign(n, 0.0);
    s.ap.assign(n, 1.0);
    s.dt.assign(n, 0.001);
    s.f_mdot.assign(F, 0.0);
    s.f_rho.assign(F, 0.0);
    s.f_qy.assign(F, 1.0);
    s.f_stress.assign(3 * (size_t)F, 0.0);
    for (auto* v : {&s.b_T, &s.b_p, &s.b_p_out, &s.b_rho, &s.b_mdot, &s.b_h}) v->assign(B, 0.0);
    s.b_qy.assign(B, 1.0);
    s.b_u.assign(3 * (size_t)B, 0.0);
    s.b_super.assign(B, 0);
    hipStream_t st = fl->solver->stream;
    up_f(fl, fl->d_f_stress, s.f_stress, st);
    up_c(fl, fl->d_ap, s.ap, st);
    up_c(fl, fl->d_ap_vec, s.ap_vec, st);
    for (auto* d : {&fl->d_u, &fl->d_u_old, &fl->d_gp}) d->alloc(3 * (size_t)n);
    for (auto* d : {&fl->d_rho, &fl->d_rho_old, &fl->d_p, &fl->d_p_old, &fl->d_pc, &fl->d_T, &fl->d_h, &fl->d_h_old,
                    &fl->d_dt})
        d->alloc(n);
    for (auto* d : {&fl->d_f_mdot, &fl->d_f_rho, &fl->d_f_qy}) d->alloc(F);
    for (auto* d : {&fl->d_b_T, &fl->d_b_rho, &fl->d_b_mdot, &fl->d_b_h, &fl->d_b_qy}) d->alloc(B);
    fl->d_b_u.alloc(3 * (size_t)B);
    fl->d_b_super.alloc(B);
    fl->d_gpc.alloc(3 * (size_t)n);
    fl->d_mach.alloc(n);
    fl->d_b_p.alloc(B);
    fl->d_b_p_out.alloc(B);
    if (fl->dev) upload_all(fl);
}

template <class T>
void copy_in(std::vector<T>& dst, const T* src) {
    if (src) std::memcpy(dst.data(), src, sizeof(T) * dst.size());
}

template <class T>
void copy_out(T* dst, const std::vector<T>& src) {
    if (dst) std::memcpy(dst, src.data(), sizeof(T) * src.size());
}

}  // namespace

extern "C" {

void ocsl_flow_default_params(ocsl_flow_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->viscous = 0;
    p->blended = 1;
    p->transient = 1;
    p->dt = 1e-3;
    p->implicit_underrelax = 0.4;
    p->implicit_underrelax_max = 0.8;
    p->ramp_iterations = 100;
    p->gamma = 1.4;
    p->molar_mass = 28.9651159 / 1000.0;
    p->t_ref = 298.15;
    p->h_formation = 0.0;
    p->sutherland[0] = 1.458e-6;
    p->sutherland[1] = 110.4;
    p->prandtl = 0.713;
    p->method_momentum = p->method_pressure = p->method_enthalpy = OCSL_ILU_BICGSTAB;
    p->mode = OCSL_MODE_REFERENCE;
    p->precision = OCSL_FP64;
    ocsl_default_options(&p->solve_options);
    p->momentum_together = 1;
}

ocsl_status ocsl_flow_create(ocsl_flow* out, const ocsl_mesh* mesh, const uint64_t* row_ptr,
                             const uint64_t* column_index, const ocsl_flow_params* prm) {
    if (!out || !mesh || !row_ptr || !column_index || !prm) return OCSL_ERR_ARG;
    *out = nullptr;
    ocsl_flow_s* fl = new (std::nothrow) ocsl_flow_s();
    if (!fl) return OCSL_ERR_NOMEM;
    ocsl_status st = guarded_f([&] {
        OCSL_REQUIRE(mesh->n_cells > 0 && mesh->n_cells < (1 << 30) && mesh->n_faces >= 0 && mesh->n_bc >= 0,
                     OCSL_ERR_ARG, "bad mesh sizes");
        OCSL_REQUIRE(prm->n_bc >= 0 && prm->n_bc <= OCSL_FLOW_MAX_BC, OCSL_ERR_ARG, "bad boundary-condition count");
        OCSL_REQUIRE(prm->transient || prm->implicit_underrelax > 0.0, OCSL_ERR_ARG, "bad under-relaxation");
        OCSL_REQUIRE(prm->support == OCSL_SUPPORT_DEVICE || prm->support == OCSL_SUPPORT_HOST, OCSL_ERR_ARG,
                     "bad support location");
        fl->prm = *prm;
        fl->dev = prm->support == OCSL_SUPPORT_DEVICE;
        const ocsl_status cs = ocsl_create(&fl->solver, mesh->n_cells, row_ptr, column_index);
        if (cs != OCSL_OK) throw Error{cs, std::string("flow solver handle: ") + ocsl_last_error()};
        HostMesh& m = fl->mesh;
        m.n = (int)mesh->n_cells;
        m.F = (int)mesh->n_faces;
        m.B = (int)mesh->n_bc;
        auto cp = [](std::vector<double>& v, const double* src, size_t cnt) {
            OCSL_REQUIRE(src || cnt == 0, OCSL_ERR_ARG, "null mesh array");
            v.assign(src, src + cnt);
        };
        auto cpi = [](std::vector<int>& v, const int32_t* src, size_t cnt) {
            OCSL_REQUIRE(src || cnt == 0, OCSL_ERR_ARG, "null mesh array");
            v.assign(src, src + cnt);
        };
        const size_t n = m.n, F = m.F, B = m.B;
        cp(m.vol, mesh->cell_volume, n);
        cpi(m.fL, mesh->face_left, F);
        cpi(m.fR, mesh->face_right, F);
        cp(m.fS, mesh->face_area_vector, 3 * F);
        cp(m.fe, mesh->face_unit, 3 * F);
        cp(m.falpha, mesh->face_alpha, 3 * F);
        cp(m.farea, mesh->face_area, F);
        cp(m.fdist, mesh->face_dist, F);
        cp(m.fksi, mesh->face_ksi, F);
        cp(m.fgeo, mesh->face_geo_aver, F);
        cpi(m.bL, mesh->bc_left, B);
        cpi(m.bmark, mesh->bc_marker, B);
        cp(m.bS, mesh->bc_area_vector, 3 * B);
        cp(m.be, mesh->bc_unit, 3 * B);
        cp(m.balpha, mesh->bc_alpha, 3 * B);
        cp(m.barea, mesh->bc_area, B);
        cp(m.bdist, mesh->bc_dist, B);
        cp(m.bksi, mesh->bc_ksi, B);
        cp(m.bperp, mesh->bc_perp, B);
        for (size_t f = 0; f < F; ++f)
            OCSL_REQUIRE(m.fL[f] >= 0 && m.fL[f] < m.n && m.fR[f] >= 0 && m.fR[f] < m.n && m.fL[f] != m.fR[f],
                         OCSL_ERR_STRUCTURE, "face cell index out of range");
        for (size_t b = 0; b < B; ++b) {
            OCSL_REQUIRE(m.bL[b] >= 0 && m.bL[b] < m.n, OCSL_ERR_STRUCTURE, "boundary face cell out of range");
            OCSL_REQUIRE(m.bmark[b] >= 0 && m.bmark[b] < prm->n_bc, OCSL_ERR_ARG, "boundary marker without condition");
        }
        // row layout: diagonal and neighbour positions in the caller's rows
        const uint64_t nnz = row_ptr[n];
        OCSL_REQUIRE(nnz < (1ull << 31), OCSL_ERR_ARG, "too many entries");
        fl->nnz_in = (int64_t)nnz;
        std::vector<int> rp32(n + 1), dpos(n, -1);
        auto find = [&](int row, int c) -> int {
            for (uint64_t q = row_ptr[row]; q < row_ptr[row + 1]; ++q)
                if (column_index[q] == (uint64_t)c) return (int)(q - row_ptr[row]);
            return -1;
        };
        for (size_t i = 0; i <= n; ++i) rp32[i] = (int)row_ptr[i];
        for (size_t i = 0; i < n; ++i) {
            OCSL_REQUIRE(row_ptr[i + 1] - row_ptr[i] <= (uint64_t)kMaxRow, OCSL_ERR_STRUCTURE,
                         "row longer than the assembly supports (16 entries)");
            dpos[i] = find((int)i, (int)i);
            OCSL_REQUIRE(dpos[i] >= 0, OCSL_ERR_STRUCTURE, "row without diagonal");
        }
        // the device's cell order (ocsl_flow_cell_order)
        OCSL_REQUIRE(prm->cell_order >= OCSL_CELL_ORDER_AUTO && prm->cell_order <= OCSL_CELL_ORDER_CURVE,
                     OCSL_ERR_ARG, "bad cell order");
        fl->cperm.resize(n);
        std::iota(fl->cperm.begin(), fl->cperm.end(), 0);
        fl->cell_perm = false;
        if (prm->cell_order != OCSL_CELL_ORDER_CALLER && n > 1) {
            // new -> old
            std::vector<int> ord = prm->cell_order == OCSL_CELL_ORDER_LOCALITY
                                       ? rcm_order(fl->solver->host)
                                       : curve_order((int)n, (int)F, m.fL.data(), m.fR.data(), m.fe.data(),
                                                     m.fdist.data());
            bool use = prm->cell_order != OCSL_CELL_ORDER_AUTO;
            if (!use) {
                // mean distance of face neighbours in each numbering
                std::vector<int> inv(n);
                for (size_t k = 0; k < n; ++k) inv[ord[k]] = (int)k;
                double dc = 0.0, dr = 0.0;
                for (size_t f = 0; f < F; ++f) {
                    dc += std::abs(m.fL[f] - m.fR[f]);
                    dr += std::abs(inv[m.fL[f]] - inv[m.fR[f]]);
                }
                use = dc > 2.0 * dr;
            }
            if (use) {
                fl->cperm = std::move(ord);
                fl->cell_perm = true;
            }
        }
        const std::vector<int>& cperm = fl->cperm;
        std::vector<int>& ciperm = fl->ciperm;
        ciperm.resize(n);
        for (size_t k = 0; k < n; ++k) ciperm[cperm[k]] = (int)k;
        if (fl->cell_perm) {
            fl->asm_lds = sizeof(double) * (size_t)(kBlock + 1) * kMaxRow;  // one padded column of slots per thread
        } else {
            int span = 1;  // at most kBlock * kMaxRow entries: 32 KB
            for (size_t i0 = 0; i0 < n; i0 += kBlock)
                span = std::max(span, rp32[std::min(n, i0 + kBlock)] - rp32[i0]);
            fl->asm_lds = sizeof(double) * (size_t)span;
        }
        fl->fperm.resize(F);
        std::iota(fl->fperm.begin(), fl->fperm.end(), 0);
        std::stable_sort(fl->fperm.begin(), fl->fperm.end(),
                         [&](int x, int y) { return ciperm[m.fL[x]] < ciperm[m.fL[y]]; });
        up(fl->d_fperm, fl->fperm, fl->solver->stream);
        OCSL_HIP(hipStreamSynchronize(fl->solver->stream));
        std::vector<int> finv(F);
        for (size_t k = 0; k < F; ++k) finv[fl->fperm[k]] = (int)k;
        // each cell's faces in the caller's ascending face order (the order
        // the reference's face loop adds them in), coded with device indices;
        // cells, neighbours and faces in the device order
        std::vector<std::vector<int>> cf(n), cb(n);
        for (size_t f = 0; f < F; ++f) {
            cf[m.fL[f]].push_back((int)f << 1);
            cf[m.fR[f]].push_back(((int)f << 1) | 1);
        }
        for (size_t b = 0; b < B; ++b) cb[m.bL[b]].push_back((int)b);
        // code = f << 5 | pos << 1 | side: faces below 2^26, positions below kMaxRow = 16
        static_assert(kMaxRow == 16, "the cell-face code holds the slot position in 4 bits");
        OCSL_REQUIRE(F < ((size_t)1 << 26), OCSL_ERR_ARG, "more than 2^26 interior faces");
        std::vector<int> cf_ptr(n + 1, 0), cf_code, cf_nb, cb_ptr(n + 1, 0), cb_face;
        for (size_t k = 0; k < n; ++k) {
            const int c = cperm[k];
            for (int code : cf[c]) {
                const int f = code >> 1;
                const int other = (code & 1) ? m.fL[f] : m.fR[f];
                const int pos = find(c, other);
                OCSL_REQUIRE(pos >= 0 && pos < kMaxRow, OCSL_ERR_STRUCTURE, "face neighbour missing from the matrix row");
                cf_code.push_back((finv[f] << 5) | (pos << 1) | (code & 1));
                cf_nb.push_back(ciperm[other]);
            }
            cf_ptr[k + 1] = (int)cf_code.size();
            for (int b : cb[c]) cb_face.push_back(b);
            cb_ptr[k + 1] = (int)cb_face.size();
        }
        // the sliced form of the cell-face lists (MeshDev::cf_sp)
        const size_t nsl = (n + 63) / 64;
        std::vector<int> cf_sp(nsl + 1, 0);
        for (size_t sl = 0; sl < nsl; ++sl) {
            int w = 0;
            for (size_t i = sl * 64; i < std::min(n, sl * 64 + 64); ++i) w = std::max(w, cf_ptr[i + 1] - cf_ptr[i]);
            OCSL_REQUIRE((int64_t)cf_sp[sl] + 64 * (int64_t)w < (1ll << 31), OCSL_ERR_ARG, "mesh too large");
            cf_sp[sl + 1] = cf_sp[sl] + 64 * w;
        }
        {
            const size_t tot = (size_t)cf_sp[nsl];
            std::vector<int> sc(tot, -1), sn(tot, 0);
            for (size_t i = 0; i < n; ++i)
                for (int k = 0; k < cf_ptr[i + 1] - cf_ptr[i]; ++k) {
                    const size_t e = (size_t)cf_sp[i / 64] + 64 * (size_t)k + i % 64, q = (size_t)cf_ptr[i] + k;
                    sc[e] = cf_code[q];
                    sn[e] = cf_nb[q];
                }
            cf_code.swap(sc);
            cf_nb.swap(sn);
        }
        hipStream_t s = fl->solver->stream;
        // cell-indexed mesh data and cell indices in the device order
        std::vector<double> vol_d(n);
        std::vector<int> dpos_d(n), rstart(n), rlen(n), fL_d(F), fR_d(F), bL_d(B);
        for (size_t k = 0; k < n; ++k) {
            const int c = cperm[k];
            vol_d[k] = m.vol[c];
            dpos_d[k] = dpos[c];
            rstart[k] = rp32[c];
            rlen[k] = rp32[c + 1] - rp32[c];
        }
        for (size_t f = 0; f < F; ++f) {
            fL_d[f] = ciperm[m.fL[f]];
            fR_d[f] = ciperm[m.fR[f]];
        }
        for (size_t b = 0; b < B; ++b) bL_d[b] = ciperm[m.bL[b]];
        up(fl->d_cperm, cperm, s);
        up(fl->d_rstart, rstart, s);
        up(fl->d_rlen, rlen, s);
        up(fl->d_vol, vol_d, s);
        up_mesh_f(fl, fl->d_fL, fL_d, s);
        up_mesh_f(fl, fl->d_fR, fR_d, s);
        up_mesh_f(fl, fl->d_fS, m.fS, s);
        up_mesh_f(fl, fl->d_fe, m.fe, s);
        up_mesh_f(fl, fl->d_fdist, m.fdist, s);
        up_mesh_f(fl, fl->d_fksi, m.fksi, s);
        up_mesh_f(fl, fl->d_fgeo, m.fgeo, s);
        up_mesh_f(fl, fl->d_falpha, m.falpha, s);
        up(fl->d_be, m.be, s);
        up(fl->d_balpha, m.balpha, s);
        up(fl->d_bdist, m.bdist, s);
        up(fl->d_bL, bL_d, s);
        up(fl->d_bmark, m.bmark, s);
        up(fl->d_bS, m.bS, s);
        up(fl->d_barea, m.barea, s);
        up(fl->d_bksi, m.bksi, s);
        up(fl->d_bperp, m.bperp, s);
        up(fl->d_cf_sp, cf_sp, s);
        up(fl->d_cf_code, cf_code, s);
        up(fl->d_cf_nb, cf_nb, s);
        up(fl->d_cb_ptr, cb_ptr, s);
        up(fl->d_cb_face, cb_face, s);
        up(fl->d_row_ptr, rp32, s);
        up(fl->d_dpos, dpos_d, s);
        fl->d_val.alloc(nnz);
        fl->d_b.alloc(n);
        fl->d_x0.alloc(n);
        // gas constants (Cfluid_properties::calc_R_specific / calc_CP_fixed)
        Gas& g = fl->gas;
        g.R = 8.31446261815324 / prm->molar_mass;
        g.gamma = prm->gamma;
        g.cp = (g.gamma * g.R) / (g.gamma - 1.0);
        g.t_ref = prm->t_ref;
        g.h_f = prm->h_formation;
        g.s0 = prm->sutherland[0];
        g.s1 = prm->sutherland[1];
        g.pr = prm->prandtl;
        fl->urf = prm->transient ? 0.5 : prm->implicit_underrelax;
        alloc_state(fl);
        OCSL_HIP(hipStreamSynchronize(s));
    });
    if (st != OCSL_OK) {
        if (fl->solver) ocsl_destroy(fl->solver);
        fl->solver = nullptr;
        delete fl;
        return st;
    }
    *out = fl;
    return OCSL_OK;
}

ocsl_status ocsl_flow_destroy(ocsl_flow fl) {
    if (!fl) return OCSL_ERR_ARG;
    ocsl_solver* h = fl->solver;
    if (h) (void)hipStreamSynchronize(h->stream);
    fl->solver = nullptr;
    delete fl;  // device buffers first, then the handle (its stream)
    if (h) ocsl_destroy(h);
    return OCSL_OK;
}

ocsl_handle ocsl_flow_solver(ocsl_flow fl) { return fl ? fl->solver : nullptr; }

ocsl_status ocsl_flow_initialize(ocsl_flow fl) {
    if (!fl) return OCSL_ERR_ARG;
    return guarded_f([&] { initialize(fl); });
}

ocsl_status ocsl_flow_set_state(ocsl_flow fl, const ocsl_flow_state* st) {
    if (!fl || !st) return OCSL_ERR_ARG;
    return guarded_f([&] {
        HostFlow& s = fl->st;
        if (fl->dev) download_all(fl);  // fields not given keep their device values
        copy_in(s.u, st->u);
        copy_in(s.u_old, st->u_old);
        copy_in(s.rho, st->rho);
        copy_in(s.rho_old, st->rho_old);
        copy_in(s.p, st->p);
        copy_in(s.p_old, st->p_old);
        copy_in(s.pc, st->p_corr);
        copy_in(s.gp, st->grad_p);
        copy_in(s.gpc, st->grad_pc);
        copy_in(s.T, st->T);
        copy_in(s.h, st->h);
        copy_in(s.h_old, st->h_old);
        copy_in(s.ap, st->ap);
        copy_in(s.ap_vec, st->ap_vec);
        copy_in(s.dt, st->dt);
        copy_in(s.mach, st->mach);
        copy_in(s.f_mdot, st->f_mdot);
        copy_in(s.f_rho, st->f_rho);
        copy_in(s.f_qy, st->f_qy);
        copy_in(s.f_stress, st->f_stress);
        copy_in(s.b_T, st->b_T);
        copy_in(s.b_p, st->b_p);
        copy_in(s.b_p_out, st->b_p_out);
        copy_in(s.b_rho, st->b_rho);
        copy_in(s.b_mdot, st->b_mdot);
        copy_in(s.b_h, st->b_h);
        copy_in(s.b_qy, st->b_qy);
        copy_in(s.b_u, st->b_u);
        copy_in(s.b_super, st->b_supersonic);
        if (fl->dev) upload_all(fl);
        else if (st->f_stress) up_f(fl, fl->d_f_stress, s.f_stress, fl->solver->stream);
        OCSL_HIP(hipStreamSynchronize(fl->solver->stream));
        fl->initialized = true;
    });
}

ocsl_status ocsl_flow_get_state(ocsl_flow fl, ocsl_flow_state* st) {
    if (!fl || !st) return OCSL_ERR_ARG;
    if (fl->dev) {
        const ocsl_status rs = guarded_f([&] { download_all(fl); });
        if (rs != OCSL_OK) return rs;
    }
    const HostFlow& s = fl->st;
    copy_out(st->u, s.u);
    copy_out(st->u_old, s.u_old);
    copy_out(st->rho, s.rho);
    copy_out(st->rho_old, s.rho_old);
    copy_out(st->p, s.p);
    copy_out(st->p_old, s.p_old);
    copy_out(st->p_corr, s.pc);
    copy_out(st->grad_p, s.gp);
    copy_out(st->grad_pc, s.gpc);
    copy_out(st->T, s.T);
    copy_out(st->h, s.h);
    copy_out(st->h_old, s.h_old);
    copy_out(st->ap, s.ap);
    copy_out(st->ap_vec, s.ap_vec);
    copy_out(st->dt, s.dt);
    copy_out(st->mach, s.mach);
    copy_out(st->f_mdot, s.f_mdot);
    copy_out(st->f_rho, s.f_rho);
    copy_out(st->f_qy, s.f_qy);
    copy_out(st->f_stress, s.f_stress);
    copy_out(st->b_T, s.b_T);
    copy_out(st->b_p, s.b_p);
    copy_out(st->b_p_out, s.b_p_out);
    copy_out(st->b_rho, s.b_rho);
    copy_out(st->b_mdot, s.b_mdot);
    copy_out(st->b_h, s.b_h);
    copy_out(st->b_qy, s.b_qy);
    copy_out(st->b_u, s.b_u);
    copy_out(st->b_supersonic, s.b_super);
    return OCSL_OK;
}

ocsl_status ocsl_flow_assemble(ocsl_flow fl, int equation, double* val, double* b, double* x0, ocsl_where where) {
    if (!fl) return OCSL_ERR_ARG;
    return guarded_f([&] {
        OCSL_REQUIRE(where == OCSL_HOST || where == OCSL_DEVICE, OCSL_ERR_ARG, "bad ocsl_where");
        assemble(fl, equation);
        hipStream_t st = fl->solver->stream;
        const auto kind = where == OCSL_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        if (val) OCSL_HIP(hipMemcpyAsync(val, fl->d_val.p, sizeof(double) * fl->nnz_in, kind, st));
        if (b) OCSL_HIP(hipMemcpyAsync(b, fl->d_b.p, sizeof(double) * fl->mesh.n, kind, st));
        if (x0) OCSL_HIP(hipMemcpyAsync(x0, fl->d_x0.p, sizeof(double) * fl->mesh.n, kind, st));
        OCSL_HIP(hipStreamSynchronize(st));
    });
}

// Algorithmic bytes of one assembly launch: every array the kernel needs,
// read or written once (face data once per face, not per side).
static double assemble_bytes(const ocsl_flow_s* fl, int eq) {
    const double n = fl->mesh.n, F = fl->mesh.F, B = fl->mesh.B, nnz = (double)fl->nnz_in;
    const ocsl_flow_params& p = fl->prm;
    double by = nnz * 8.0 + n * (8.0 + 8.0 + 4.0 * 4.0 + 8.0);  // values, b, x0, row maps, volume
    by += 2.0 * F * 8.0 + F * (8.0 + 8.0 + 8.0);                 // cell-face lists; L, R, geo, mass flux
    by += B * (4.0 + 4.0 + 8.0);                                  // boundary list, marker, mass flux
    if (p.transient) by += n * 8.0;                               // dt
    if (eq <= OCSL_EQ_MOMENTUM_Z) {
        by += n * (8.0 + 8.0 + 8.0 + 8.0 + 8.0 + 16.0 + 8.0);     // u_d, u_old_d, rho, rho_old, dp/dx_d, ap rw, ap_vec
        by += B * 24.0;                                           // boundary velocity
        if (p.viscous) by += n * 8.0 + F * (8.0 + 24.0) + B * (24.0 + 16.0);  // T, ksi, stress; walls
    } else if (eq == OCSL_EQ_PRESSURE) {
        by += n * (8.0 * 4.0) + F * 16.0 + B * 28.0;              // T, rho, rho_old, p'; rho_f, Qy; boundary
    } else {
        by += n * (8.0 * 7.0) + F * 8.0 + B * 16.0;               // h, h_old, T, rho, rho_old, p, p_old; ksi
    }
    return by;
}

ocsl_status ocsl_flow_time_assemble(ocsl_flow fl, int equation, int reps, double* avg_ms, double* bytes) {
    if (!fl || !avg_ms || !bytes || reps <= 0) return OCSL_ERR_ARG;
    return guarded_f([&] {
        OCSL_REQUIRE(equation >= OCSL_EQ_MOMENTUM_X && equation <= OCSL_EQ_ENTHALPY, OCSL_ERR_ARG, "bad equation");
        ocsl_solver* h = fl->solver;
        if (!fl->dev) upload_cells(fl, equation);
        DBuf<double> ap_saved;
        if (fl->dev) {  // the repeated momentum launches accumulate into ap
            ap_saved.alloc(fl->mesh.n);
            OCSL_HIP(hipMemcpyAsync(ap_saved.p, fl->d_ap.p, sizeof(double) * fl->mesh.n, hipMemcpyDeviceToDevice,
                                    h->stream));
        }
        auto launch = [&] {
            if (equation <= OCSL_EQ_MOMENTUM_Z) launch_assemble<kMomentum>(fl, equation);
            else if (equation == OCSL_EQ_PRESSURE) launch_assemble<kPressure>(fl, 0);
            else launch_assemble<kEnthalpy>(fl, 0);
        };
        launch();  // warm
        OCSL_HIP(hipEventRecord(h->ev0, h->stream));
        for (int r = 0; r < reps; ++r) launch();
        OCSL_HIP(hipEventRecord(h->ev1, h->stream));
        OCSL_HIP(hipStreamSynchronize(h->stream));
        float ms = 0.f;
        OCSL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *avg_ms = ms / reps;
        *bytes = assemble_bytes(fl, equation);
        // the repeated momentum launches accumulated into ap: restore it
        if (fl->dev)
            OCSL_HIP(hipMemcpyAsync(fl->d_ap.p, ap_saved.p, sizeof(double) * fl->mesh.n, hipMemcpyDeviceToDevice,
                                    h->stream));
        else
            up_c(fl, fl->d_ap, fl->st.ap, h->stream);
        OCSL_HIP(hipStreamSynchronize(h->stream));
    });
}

ocsl_status ocsl_flow_viscous_stress(ocsl_flow fl) {
    if (!fl) return OCSL_ERR_ARG;
    return guarded_f([&] { viscous_stress(fl); });
}

ocsl_status ocsl_flow_iterate(ocsl_flow fl, int end_of_step) {
    if (!fl) return OCSL_ERR_ARG;
    return guarded_f([&] {
        OCSL_REQUIRE(fl->initialized, OCSL_ERR_STATE, "flow not initialized");
        iterate(fl);
        if (end_of_step) update_timestep(fl);
    });
}

ocsl_status ocsl_flow_run(ocsl_flow fl, int steps, int inner) {
    if (!fl || steps < 0 || inner < 1) return OCSL_ERR_ARG;
    return guarded_f([&] {
        OCSL_REQUIRE(fl->initialized, OCSL_ERR_STATE, "flow not initialized");
        const int in = fl->prm.transient ? inner : 1;  // Croot::initialize_solver (croot.cpp:92-93)
        for (int s = 0; s < steps; ++s) {
            for (int k = 0; k < in; ++k) iterate(fl);
            update_timestep(fl);
        }
    });
}

ocsl_status ocsl_flow_cell_order_get(ocsl_flow fl, int32_t* order) {
    if (!fl || !order) return OCSL_ERR_ARG;
    std::copy(fl->cperm.begin(), fl->cperm.end(), order);
    return OCSL_OK;
}

ocsl_status ocsl_flow_solve_history(ocsl_flow fl, double* out, int64_t cap, int64_t* len) {
    if (!fl || !len) return OCSL_ERR_ARG;
    *len = (int64_t)fl->hist.size();
    if (out)
        for (int64_t k = 0; k < std::min<int64_t>(cap, *len); ++k) out[k] = fl->hist[k];
    return OCSL_OK;
}

}  // extern "C"
