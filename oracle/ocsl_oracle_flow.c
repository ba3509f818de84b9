/*
 * ocsl_oracle_flow.c -- CPU restatement of the reference's coefficient
 * assembly (cflow_solver_cell_face_loops.cpp with the stencil helpers of
 * solvers_common.h and the gas model of cfluid_properties.cpp).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the GPU assembly kernels and,
 * timed, the CPU baseline of that loop.  The product path never links it.
 *
 * It follows the reference's structure: the matrix and b are cleared, then the
 * face loop, the boundary-face loop and the cell loop add their terms into
 * the CSR slots found by a linear search of the row (add_insert_element,
 * cmatrix_solver.cpp:105-152, with `omp atomic` when threaded), in face
 * order.  Inputs are the structs of include/ocsl_flow.h (host arrays).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "ocsl_flow.h"

extern int or_get_threads(void);

static const double kAlpha = 0.7; /* solvers_common.h:30 */

typedef struct {
    double R, gamma, cp, t_ref, h_f, s0, s1, pr;
} gas_t;

static gas_t gas_of(const ocsl_flow_params* p) {
    gas_t g;
    g.R = 8.31446261815324 / p->molar_mass;          /* calc_R_specific */
    g.gamma = p->gamma;
    g.cp = (g.gamma * g.R) / (g.gamma - 1.0);        /* calc_CP_fixed */
    g.t_ref = p->t_ref;
    g.h_f = p->h_formation;
    g.s0 = p->sutherland[0];
    g.s1 = p->sutherland[1];
    g.pr = p->prandtl;
    return g;
}
/* cfluid_properties.cpp:68-75, 122-128 */
static double dp_drho(const gas_t* g, double T) {
    double r = g->R * T;
    r = r * g->gamma;
    return 1.0 / r;
}
static double viscosity(const gas_t* g, double T) {
    double r = g->s0 * pow(T, 3.0 / 2.0);
    return r / (T + g->s1);
}

/* slot of (row, col) in the caller's CSR, -1 if absent (CCSR::get_element) */
static int64_t slot(const uint64_t* rp, const uint64_t* col, int64_t row, int64_t c) {
    for (uint64_t q = rp[row]; q < rp[row + 1]; ++q)
        if (col[q] == (uint64_t)c) return (int64_t)q;
    return -1;
}

static void add_a(double* val, const uint64_t* rp, const uint64_t* col, int64_t i, int64_t j, double v) {
    const int64_t q = slot(rp, col, i, j);
    if (q < 0) return;
#pragma omp atomic
    val[q] += v;
}
static void add_b(double* b, int64_t i, double v) {
#pragma omp atomic
    b[i] += v;
}

static double face_avg(double geo, double l, double r) {
    double v = geo * l;
    v += (1.0 - geo) * r;
    return v;
}
/* Convective_face_interp (solvers_common.h:64-88) */
static double conv_interp(int blend, double geo, double mdot, double l, double r) {
    const double upw = 1.0 - kAlpha;
    if (mdot >= 0.0) return blend ? l * upw + kAlpha * face_avg(geo, l, r) : l;
    return blend ? r * upw + kAlpha * face_avg(geo, l, r) : r;
}
/* Convective_flux (solvers_common.h:90-135) */
static void conv_flux(double* val, double* b, const uint64_t* rp, const uint64_t* col, int blend, int64_t L,
                      int64_t R, double geo, double mdot, double lv, double rv) {
    if (mdot >= 0.0) {
        add_a(val, rp, col, L, L, mdot);
        add_a(val, rp, col, R, L, -mdot);
        if (blend) {
            const double v = kAlpha * (face_avg(geo, lv, rv) - lv);
            add_b(b, L, -mdot * v);
            add_b(b, R, mdot * v);
        }
    } else if (mdot < 0.0) {
        add_a(val, rp, col, L, R, mdot);
        add_a(val, rp, col, R, R, -mdot);
        if (blend) {
            const double v = kAlpha * (face_avg(geo, lv, rv) - rv);
            add_b(b, L, -mdot * v);
            add_b(b, R, mdot * v);
        }
    }
}
/* Diffusive_flux (solvers_common.h:42-53) */
static void diff_flux(double* val, const uint64_t* rp, const uint64_t* col, int64_t L, int64_t R, double D) {
    add_a(val, rp, col, L, L, D);
    add_a(val, rp, col, L, R, -D);
    add_a(val, rp, col, R, R, D);
    add_a(val, rp, col, R, L, -D);
}

/*
 * System of equation `eq` (0..2 momentum of that dimension, 3 pressure
 * correction, 4 enthalpy) from the state.  For momentum, ap[i] += a_ii / 3
 * and ap_vec[3 i + d] = a_ii as the reference's cell loop does.  urf is the
 * implicit under-relaxation (steady runs).  Returns 0, or -1 on bad input.
 */
int or_assemble(int eq, const ocsl_mesh* m, const ocsl_flow_state* s, const ocsl_flow_params* p, double urf,
                const uint64_t* rp, const uint64_t* col, double* val, double* b, double* x0) {
    if (eq < 0 || eq > 4 || !m || !s || !p) return -1;
    const gas_t g = gas_of(p);
    const int64_t n = m->n_cells, F = m->n_faces, B = m->n_bc;
    const int d = eq < 3 ? eq : 0;
    const int blend = p->blended, visc = p->viscous, trans = p->transient;
    memset(val, 0, sizeof(double) * rp[n]);
    memset(b, 0, sizeof(double) * (size_t)n);
    const int th = or_get_threads();
    /* face loop (cflow_solver.cpp:268-272, 323-327, 429-433) */
#pragma omp parallel for num_threads(th) schedule(static)
    for (int64_t f = 0; f < F; ++f) {
        const int64_t L = m->face_left[f], R = m->face_right[f];
        const double geo = m->face_geo_aver[f], mdot = s->f_mdot[f];
        if (eq < 3) { /* calc_flux_momentum, loops.cpp:38-58 */
            conv_flux(val, b, rp, col, blend, L, R, geo, mdot, s->u[3 * L + d], s->u[3 * R + d]);
            if (visc) {
                const double Tf = face_avg(geo, s->T[L], s->T[R]);
                double D = viscosity(&g, Tf);
                D = D * m->face_ksi[f];
                diff_flux(val, rp, col, L, R, D);
                add_b(b, L, s->f_stress[3 * f + d]);
                add_b(b, R, -s->f_stress[3 * f + d]);
            }
        } else if (eq == 3) { /* calc_flux_pressure, loops.cpp:186-218 */
            add_b(b, L, -mdot);
            add_b(b, R, mdot);
            const double Tf = conv_interp(blend, geo, mdot, s->T[L], s->T[R]);
            double dum = dp_drho(&g, Tf);
            dum = dum / s->f_rho[f];
            dum *= mdot;
            conv_flux(val, b, rp, col, blend, L, R, geo, dum, s->p_corr[L], s->p_corr[R]);
            diff_flux(val, rp, col, L, R, s->f_qy[f] * s->f_rho[f]);
        } else { /* calc_flux_enthalpy, loops.cpp:334-346 */
            conv_flux(val, b, rp, col, blend, L, R, geo, mdot, s->h[L], s->h[R]);
            const double Tf = face_avg(geo, s->T[L], s->T[R]);
            double D = viscosity(&g, Tf);
            D = D / g.pr * m->face_ksi[f];
            diff_flux(val, rp, col, L, R, D);
        }
    }
    /* boundary-face loop */
#pragma omp parallel for num_threads(th) schedule(static)
    for (int64_t bf = 0; bf < B; ++bf) {
        const int64_t i = m->bc_left[bf];
        const ocsl_bc* bc = &p->bc[m->bc_marker[bf]];
        const double bm = s->b_mdot[bf];
        if (eq < 3) { /* loops.cpp:77-132 */
            if (bc->type == OCSL_BC_WALL && visc && bc->no_slip) {
                const double inv = 1.0 / m->bc_area[bf];
                double nrm[3];
                for (int k = 0; k < 3; ++k) nrm[k] = inv * m->bc_area_vector[3 * bf + k];
                double D = viscosity(&g, s->T[i]);
                D *= m->bc_area[bf] / m->bc_perp[bf];
                add_a(val, rp, col, i, i, D * (1.0 - (nrm[d] * nrm[d])));
                double dum = 0.0;
                for (int k = d + 1; k < d + 3; ++k) {
                    const int k_ = k % 3;
                    dum += s->u[3 * i + k_] * nrm[k_] * nrm[d];
                }
                add_b(b, i, D * dum);
            }
            if (bc->type == OCSL_BC_VELOCITY_INLET || bc->type == OCSL_BC_SUPERSONIC_INLET)
                add_b(b, i, -(bm * s->b_u[3 * bf + d]));
            if (bc->type == OCSL_BC_PRESSURE_OUTLET) add_a(val, rp, col, i, i, bm);
        } else if (eq == 3) { /* loops.cpp:229-297 */
            if (bc->type == OCSL_BC_VELOCITY_INLET) {
                add_b(b, i, -bm);
                double dum = dp_drho(&g, s->b_T[bf]) / s->rho[i];
                dum *= bm;
                if (dum >= 0.0 || dum < 0.0) add_a(val, rp, col, i, i, dum);
            }
            if (bc->type == OCSL_BC_SUPERSONIC_INLET) {
                add_b(b, i, -bm);
                const double D = s->b_qy[bf] * s->b_rho[bf];
                add_a(val, rp, col, i, i, D);
                add_b(b, i, D * 0.0);
            }
            if (bc->type == OCSL_BC_PRESSURE_OUTLET) {
                add_b(b, i, -bm);
                if (s->b_supersonic[bf]) {
                    double dum = dp_drho(&g, s->T[i]) / s->rho[i];
                    dum *= bm;
                    if (dum >= 0.0 || dum < 0.0) add_a(val, rp, col, i, i, dum);
                } else {
                    const double D = s->b_qy[bf] * s->rho[i];
                    add_a(val, rp, col, i, i, D);
                    add_b(b, i, D * 0.0);
                }
            }
        } else { /* loops.cpp:357-421 */
            if (bc->type == OCSL_BC_VELOCITY_INLET || bc->type == OCSL_BC_SUPERSONIC_INLET) {
                add_b(b, i, -(bm * s->b_h[bf]));
                double D = viscosity(&g, s->b_T[bf]);
                D *= g.pr * m->bc_ksi[bf];
                add_a(val, rp, col, i, i, D);
                add_b(b, i, D * s->b_h[bf]);
            }
            if (bc->type == OCSL_BC_PRESSURE_OUTLET && (bm >= 0.0 || bm < 0.0)) add_a(val, rp, col, i, i, bm);
        }
    }
    /* cell loop */
#pragma omp parallel for num_threads(th) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double V = m->cell_volume[i];
        const int64_t qd = slot(rp, col, i, i);
        if (eq < 3) { /* loops.cpp:143-175 */
            double dum = 0.0, dum2 = 0.0;
            if (trans) {
                dum = V / s->dt[i];
                dum2 = dum * s->rho[i];
                val[qd] += dum2;
                dum2 = dum * (s->u_old[3 * i + d] * s->rho_old[i]);
            }
            dum2 = dum2 - s->grad_p[3 * i + d] * V;
            b[i] += dum2;
            if (!trans) {
                const double dd = val[qd];
                val[qd] = dd / urf;
                b[i] += (1.0 - urf) / urf * dd * s->u_old[3 * i + d];
            }
            s->ap[i] += val[qd] / 3.0;
            s->ap_vec[3 * i + d] = val[qd];
            x0[i] = s->u[3 * i + d];
        } else if (eq == 3) { /* loops.cpp:299-322 */
            if (trans) {
                const double dum = V / s->dt[i];
                double dum2 = dum * (s->rho[i] - s->rho_old[i]);
                b[i] += -dum2;
                dum2 = dum * dp_drho(&g, s->T[i]);
                val[qd] += dum2;
            } else {
                val[qd] = val[qd] / urf;
            }
            x0[i] = s->p_corr[i];
        } else { /* loops.cpp:423-446 */
            if (trans) {
                const double dum = V / s->dt[i];
                double dum2 = dum * s->rho[i];
                val[qd] += dum2;
                dum2 = dum * (s->h_old[i] * s->rho_old[i]);
                dum2 = dum2 + (s->p[i] - s->p_old[i]) / s->dt[i] * V;
                b[i] += dum2;
            } else {
                const double dd = val[qd];
                val[qd] = dd / urf;
                b[i] += (1.0 - urf) / urf * dd * s->h_old[i];
            }
            x0[i] = s->h[i];
        }
    }
    return 0;
}
