"""GPU coefficient assembly and the flow solver around it, against the
reference's own runs (golden fixtures from tests/golden/make_golden.py): the
system assembled from each recorded flow state, and whole runs (fields and
per-solve residuals)."""
import os

import numpy as np
import pytest

import parity
from opencslsolver_amd import flow
from opencslsolver_amd.gen import CSR
from opencslsolver_amd.gen import hex_mesh as gen_hex

from flowcases import CASES, FLOW_META, GOLDEN, SOD_META, params_for, records, state_at

pytestmark = pytest.mark.gpu


def _check_history(name, hist, ref, limit):
    """Per-solve residuals.  Values at or below 1e-14 are rounding noise: a
    system whose right-hand side cancels to zero here (the reference's
    contracted multiply-adds leave ~1e-16 of it) gives the reference's 0/0
    residual (cmatrix_solver.cpp:733), so there both sides only have to be
    noise."""
    parity.check(f"{name} residual history", parity.hist_rel(hist, ref), limit)


def _check_fields(name, st, d, spread_keys):
    for f in ("p", "T", "rho", "h", "u"):
        parity.check(f"{name} {f}", parity.rel(st[f], d[f]), parity.bar(*spread_keys, f))


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def _close(a, b, rtol):
    return _rel(a, b) <= rtol


@pytest.mark.parametrize("case", CASES)
def test_assembly_matches_reference_states(case):
    d = np.load(os.path.join(GOLDEN, f"flow_{case}.npz"))
    cfg = FLOW_META[case]
    with flow.Flow(flow.mesh_from_npz(d), d["row_ptr"], d["col"], params_for(cfg)) as fl:
        fl.initialize()
        for k in range(records(case, d)):
            eq, st = state_at(d, k, fl.n)
            fl.set_state(**st)
            val, b, x0 = fl.assemble(eq)
            assert _close(val, d["dump_val"][k], 1e-12), (case, k)
            assert _close(b, d["dump_b"][k], 1e-12), (case, k)
            assert np.array_equal(x0, d["dump_x0"][k]), (case, k)
            if eq <= flow.MOMENTUM_Z:
                out = fl.get_state()
                assert _close(out["ap"], d["state_ap"][k], 1e-12), (case, k)
                assert _close(out["ap_vec"][:, eq], d["state_ap_vec"][k][:, eq], 1e-12), (case, k)


def test_assembly_rows_conserve_on_symmetric_diffusion():
    """Closed form: with zero mass fluxes and no cell terms missing, a viscous
    momentum row sums to its transient term rho V / dt (diffusion rows sum to
    zero), and the matrix is symmetric."""
    d = np.load(os.path.join(GOLDEN, "flow_duct64.npz"))
    cfg = dict(FLOW_META["duct64"])
    cfg["bc_type"] = [flow.WALL, flow.EMPTY, flow.WALL]
    cfg["bc_no_slip"] = [0, 0, 0]
    with flow.Flow(flow.mesh_from_npz(d), d["row_ptr"], d["col"], params_for(cfg)) as fl:
        n, F, B = fl.n, fl.F, fl.B
        rng = np.random.default_rng(5)
        T = 300 + 10 * rng.random(n)
        rho = 1 + rng.random(n)
        fl.set_state(u=rng.random((n, 3)), u_old=rng.random((n, 3)), rho=rho, rho_old=rho, T=T,
                     grad_p=np.zeros((n, 3)), dt=np.full(n, 1e-4), ap=np.zeros(n), f_mdot=np.zeros(F),
                     f_stress=np.zeros((F, 3)), b_mdot=np.zeros(B), b_u=np.zeros((B, 3)))
        val, b, _ = fl.assemble(flow.MOMENTUM_X)
    a = CSR(d["row_ptr"], d["col"], val)
    dense = a.to_dense()
    vol = d["mesh_cell_volume"]
    assert np.allclose(dense.sum(axis=1), rho * vol / 1e-4, rtol=1e-12)
    assert np.allclose(dense, dense.T, rtol=0, atol=1e-12 * np.abs(dense).max())


@pytest.mark.parametrize("case", CASES)
def test_flow_run_matches_reference(case):
    d = np.load(os.path.join(GOLDEN, f"flow_{case}.npz"))
    cfg = FLOW_META[case]
    with flow.Flow(flow.mesh_from_npz(d), d["row_ptr"], d["col"], params_for(cfg)) as fl:
        fl.initialize()
        fl.run(cfg["steps"], cfg["inner"])
        st = fl.get_state()
        hist = fl.solve_history()
    _check_fields(f"flow {case}", st, d, ("flow_runs", case))
    assert len(hist) == int(d["count"])
    _check_history(f"flow {case}", hist, d["res_hist"], parity.bar("flow_runs", case, "hist"))


@pytest.mark.parametrize("case", CASES)
def test_flow_support_device_matches_host(case):
    """The SIMPLE step's support loops as device kernels over the resident
    state (the default) against their host restatement, on the reference
    configurations (walls, inlets, outlets, supersonic, steady): the same run
    to rounding.  The device kernels contract multiply-adds and the host code
    does not; after the fixture's steps the fields differ by ~1e-11 relative,
    so their bar is the reference-parity bar of the run test above, 1e-10."""
    d = np.load(os.path.join(GOLDEN, f"flow_{case}.npz"))
    cfg = FLOW_META[case]
    out = {}
    for support in (flow.SUPPORT_DEVICE, flow.SUPPORT_HOST):
        prm = params_for(cfg)
        prm.support = support
        with flow.Flow(flow.mesh_from_npz(d), d["row_ptr"], d["col"], prm) as fl:
            fl.initialize()
            fl.run(cfg["steps"], cfg["inner"])
            out[support] = (fl.get_state(), fl.solve_history())
    dev, host = out[flow.SUPPORT_DEVICE][0], out[flow.SUPPORT_HOST][0]
    for key in ("p", "T", "rho", "h", "u", "mach", "f_rho", "b_p", "b_rho", "b_h", "b_T"):
        assert _close(dev[key], host[key], 1e-10), (case, key)
    # corrections and fluxes are small differences of the fields (p' is ~1e-3
    # of p and comes from a 4-iteration pressure solve): rounding differences
    # of the fields show at ~1e-8 of their own size
    for key in ("p_corr", "grad_p", "grad_pc", "f_mdot", "f_qy", "b_mdot", "b_qy"):
        assert _close(dev[key], host[key], 1e-6), (case, key)
    assert np.array_equal(dev["b_supersonic"], host["b_supersonic"])
    assert len(out[flow.SUPPORT_DEVICE][1]) == len(out[flow.SUPPORT_HOST][1])


@pytest.mark.parametrize("viscous,blended,transient", [(1, 1, 1), (0, 0, 0)])
def test_flow_3d_device_run_matches_host(viscous, blended, transient):
    """Two SIMPLE steps (two inner iterations each) on a permuted 8^3 hex mesh
    with a velocity inlet, a pressure outlet, no-slip and slip walls and a
    symmetry plane: device-resident support loops against the host
    restatement, TOLERANCE-mode FP64 solves."""
    rng = np.random.default_rng(5)
    n = 8
    perm = rng.permutation(n ** 3)
    mesh, rp, col = gen_hex(n, perm=perm)
    bcs = [dict(type=flow.VELOCITY_INLET, velocity=(3.0, 0.5, -0.2), temperature=300.0),
           dict(type=flow.PRESSURE_OUTLET, pressure=1e5),
           dict(type=flow.WALL, no_slip=1), dict(type=flow.WALL, no_slip=0),
           dict(type=flow.SYMMETRY), dict(type=flow.SYMMETRY)]
    out = {}
    for support in (flow.SUPPORT_DEVICE, flow.SUPPORT_HOST):
        prm = flow.default_params(viscous=viscous, blended=blended, transient=transient, dt=1e-4,
                                  implicit_underrelax=0.6, p0=1e5, t0=300.0, u0=(2.0, 0.0, 0.0), bcs=bcs,
                                  mode=1, method_pressure=0, support=support,
                                  solve_options=dict(tol=1e-12, max_iter=2000))
        with flow.Flow(mesh, rp, col, prm) as fl:
            fl.initialize()
            fl.run(2, 2)
            out[support] = fl.get_state()
    dev, host = out[flow.SUPPORT_DEVICE], out[flow.SUPPORT_HOST]
    for key in ("p", "T", "rho", "h", "u", "f_mdot", "b_mdot", "b_p", "b_rho"):
        assert np.all(np.isfinite(dev[key])), key
        assert _close(dev[key], host[key], 1e-9), key


@pytest.mark.parametrize("viscous,walls", [(0, "no_slip"), (1, "slip")])
def test_momentum_solved_together_changes_no_bit(viscous, walls):
    """The three momentum systems solved by one REFERENCE solve of three
    right-hand sides (the default where no no-slip wall of a viscous flow
    couples them) against three consecutive solves: bitwise the same
    fields and per-solve residuals, on a permuted 8^3 hex mesh."""
    rng = np.random.default_rng(9)
    n = 8
    perm = rng.permutation(n ** 3)
    mesh, rp, col = gen_hex(n, perm=perm)
    bcs = [dict(type=flow.VELOCITY_INLET, velocity=(3.0, 0.5, -0.2), temperature=300.0),
           dict(type=flow.PRESSURE_OUTLET, pressure=1e5),
           dict(type=flow.WALL, no_slip=1 if walls == "no_slip" else 0), dict(type=flow.WALL, no_slip=0),
           dict(type=flow.SYMMETRY), dict(type=flow.SYMMETRY)]
    out = {}
    for together in (0, 1):
        prm = flow.default_params(viscous=viscous, blended=1, transient=1, dt=1e-4, p0=1e5, t0=300.0,
                                  u0=(2.0, 0.0, 0.0), bcs=bcs, method_pressure=0, momentum_together=together)
        with flow.Flow(mesh, rp, col, prm) as fl:
            fl.initialize()
            fl.run(2, 2)
            out[together] = (fl.get_state(), fl.solve_history())
    (s0, h0), (s1, h1) = out[0], out[1]
    assert np.array_equal(h0, h1, equal_nan=True)
    for key in ("p", "T", "rho", "h", "u", "f_mdot", "b_mdot"):
        assert np.all(np.isfinite(s1[key])), key
        assert np.array_equal(s0[key], s1[key]), key


def test_sod_momentum_together_changes_no_bit():
    """The shock tube (serial triangular solves of three interleaved
    right-hand sides): two steps with the momentum systems solved together
    and one by one, bitwise the same."""
    d = np.load(os.path.join(GOLDEN, "sod_split_2steps.npz"))
    mesh = np.load(os.path.join(GOLDEN, "flow_sod1000_mesh.npz"))
    cfg = dict(FLOW_META["sod64"])
    cfg.update(n_cells=SOD_META["n_cells"], dt=SOD_META["dt"], inner=SOD_META["inner"], solver_type=0,
               solver_mode=1)
    out = {}
    for together in (0, 1):
        prm = params_for(cfg)
        prm.momentum_together = together
        with flow.Flow(flow.mesh_from_npz(mesh), d["row_ptr"], d["col"], prm) as fl:
            fl.initialize()
            fl.run(2, SOD_META["inner"])
            out[together] = (fl.get_state(), fl.solve_history())
    assert np.array_equal(out[0][1], out[1][1], equal_nan=True)
    for key in ("p", "T", "rho", "h", "u"):
        assert np.array_equal(out[0][0][key], out[1][0][key]), key


def _sod(tag, steps):
    d = np.load(os.path.join(GOLDEN, f"sod_{tag}_{steps}steps.npz"))
    mesh = np.load(os.path.join(GOLDEN, "flow_sod1000_mesh.npz"))
    cfg = dict(FLOW_META["sod64"])
    cfg.update(n_cells=SOD_META["n_cells"], dt=SOD_META["dt"], inner=SOD_META["inner"])
    cfg["solver_type"] = 0 if tag in ("cg", "split") else 1
    cfg["solver_mode"] = 1 if tag == "split" else 0
    assert np.array_equal(mesh["row_ptr"], d["row_ptr"]) and np.array_equal(mesh["col"], d["col"])
    with flow.Flow(flow.mesh_from_npz(mesh), d["row_ptr"], d["col"], params_for(cfg)) as fl:
        fl.initialize()
        fl.run(steps, SOD_META["inner"])
        return d, fl.get_state(), fl.solve_history()


@pytest.mark.parametrize("tag", ["bicgstab", "split"])
def test_sod_two_steps(tag):
    d, st, hist = _sod(tag, 2)
    _check_fields(f"sod {tag} 2 steps", st, d, ("sod_runs", f"{tag}_2"))
    _check_history(f"sod {tag} 2 steps", hist, d["res_hist"], parity.bar("sod_runs", f"{tag}_2", "hist"))


def test_sod_200_steps_split():
    """The shock tube to 200 steps (CG for the pressure correction, ILU(0)-
    BiCGSTAB for momentum and enthalpy): fields and per-solve residuals
    against the reference's.  Over 2000 SIMPLE iterations rounding
    differences grow: the reference's two builds (-O3 -ffast-math
    -march=native and -O2 -ffp-contract=off) end 8e-11 (p), 1.4e-10 (T),
    4.1e-10 (u) apart, and their per-solve residuals up to 5.7e-6 apart; the
    bars are five times those spreads."""
    d, st, hist = _sod("split", 200)
    _check_fields("sod split 200 steps", st, d, ("sod_runs", "split_200"))
    assert len(hist) == int(d["count"])
    _check_history("sod split 200 steps", hist, d["res_hist"], parity.bar("sod_runs", "split_200", "hist"))


def _random_state(rng, n, F, B):
    return dict(
        u=rng.normal(size=(n, 3)), u_old=rng.normal(size=(n, 3)), rho=1.0 + rng.random(n),
        rho_old=1.0 + rng.random(n), p=1e5 + 1e3 * rng.random(n), p_old=1e5 + 1e3 * rng.random(n),
        p_corr=rng.normal(size=n), grad_p=rng.normal(size=(n, 3)), T=280.0 + 40.0 * rng.random(n),
        h=3e5 + 1e4 * rng.random(n), h_old=3e5 + 1e4 * rng.random(n), dt=np.full(n, 1e-4),
        ap=rng.random(n), ap_vec=rng.random((n, 3)),
        f_mdot=rng.normal(size=F) * 1e-3, f_rho=1.0 + rng.random(F), f_qy=rng.random(F) * 1e-3,
        f_stress=rng.normal(size=(F, 3)) * 1e-5,
        b_T=280.0 + 40.0 * rng.random(B), b_p=1e5 + 1e3 * rng.random(B), b_rho=1.0 + rng.random(B),
        b_mdot=rng.normal(size=B) * 1e-3, b_h=3e5 + 1e4 * rng.random(B), b_qy=rng.random(B) * 1e-3,
        b_u=rng.normal(size=(B, 3)), b_supersonic=(rng.random(B) < 0.5).astype(np.int32))


@pytest.mark.parametrize("viscous,blended,transient", [(1, 1, 1), (1, 0, 0), (0, 1, 0), (0, 0, 1)])
def test_assembly_3d_matches_oracle(viscous, blended, transient):
    """A permuted 10^3 hex mesh with every boundary kind (velocity inlet,
    pressure outlet, no-slip and slip walls, symmetry, supersonic inlet) and a
    random state: the GPU rows against the oracle's scatter loops."""
    import oracle

    oracle.set_threads(1)
    rng = np.random.default_rng(11)
    n = 10
    perm = rng.permutation(n ** 3)
    mesh, rp, col = gen_hex(n, perm=perm)
    bcs = [dict(type=flow.VELOCITY_INLET, velocity=(3.0, 0.5, -0.2), temperature=300.0),
           dict(type=flow.PRESSURE_OUTLET, pressure=1e5),
           dict(type=flow.WALL, no_slip=1), dict(type=flow.WALL, no_slip=0),
           dict(type=flow.SYMMETRY), dict(type=flow.SUPERSONIC_INLET, velocity=(400.0, 0.0, 0.0), pressure=1e5,
                                          temperature=300.0)]
    prm = flow.default_params(viscous=viscous, blended=blended, transient=transient, implicit_underrelax=0.6,
                              bcs=bcs)
    with flow.Flow(mesh, rp, col, prm) as fl:
        st = _random_state(rng, fl.n, fl.F, fl.B)
        for eq in range(5):
            fl.set_state(**st)
            val, b, x0 = fl.assemble(eq)
            v_o, b_o, x_o, ap_o, apv_o = oracle.assemble(eq, mesh, rp, col, prm, st, urf=0.6)
            assert _close(val, v_o, 1e-12), eq
            assert _close(b, b_o, 1e-12), eq
            assert np.array_equal(x0, x_o), eq
            if eq < 3:
                out = fl.get_state()
                assert _close(out["ap"], ap_o, 1e-12), eq
                assert _close(out["ap_vec"][:, eq], apv_o[:, eq], 1e-12), eq


@pytest.mark.parametrize("support", [flow.SUPPORT_DEVICE, flow.SUPPORT_HOST])
def test_cell_order_changes_no_bit(support):
    """The device's cell order (ocsl_flow_cell_order) moves data, not
    arithmetic: on a randomly numbered 9^3 hex mesh with every boundary kind,
    two SIMPLE steps give bitwise the same fields, fluxes and per-solve
    residuals whether the device keeps the caller's order, the reverse
    Cuthill-McKee order or the Hilbert-curve order; so does one assembly of
    each equation.  AUTO picks the curve order here."""
    rng = np.random.default_rng(21)
    n = 9
    mesh, rp, col = gen_hex(n, perm=rng.permutation(n ** 3))
    bcs = [dict(type=flow.VELOCITY_INLET, velocity=(3.0, 0.5, -0.2), temperature=300.0),
           dict(type=flow.PRESSURE_OUTLET, pressure=1e5),
           dict(type=flow.WALL, no_slip=1), dict(type=flow.WALL, no_slip=0),
           dict(type=flow.SYMMETRY), dict(type=flow.SUPERSONIC_INLET, velocity=(400.0, 0.0, 0.0), pressure=1e5,
                                          temperature=300.0)]
    out = {}
    orders = (flow.CELL_ORDER_CALLER, flow.CELL_ORDER_LOCALITY, flow.CELL_ORDER_CURVE, flow.CELL_ORDER_AUTO)
    for order in orders:
        prm = flow.default_params(viscous=1, blended=1, transient=1, dt=1e-4, implicit_underrelax=0.6, p0=1e5,
                                  t0=300.0, u0=(2.0, 0.0, 0.0), bcs=bcs, mode=1, method_pressure=0, support=support,
                                  solve_options=dict(tol=1e-12, max_iter=2000), cell_order=order)
        with flow.Flow(mesh, rp, col, prm) as fl:
            co = fl.cell_order()
            fl.initialize()
            systems = [fl.assemble(eq) for eq in range(5)]
            fl.initialize()
            fl.run(2, 2)
            out[order] = (co, systems, fl.get_state(), fl.solve_history())
    caller, local, curve, auto = (out[o] for o in orders)
    assert np.array_equal(caller[0], np.arange(n ** 3))
    for other in (local, curve):
        assert not np.array_equal(other[0], np.arange(n ** 3))
        assert np.array_equal(np.sort(other[0]), np.arange(n ** 3))
    assert not np.array_equal(local[0], curve[0])
    assert np.array_equal(auto[0], curve[0])
    for other in (local, curve):
        for eq in range(5):
            for k in range(3):
                assert np.array_equal(caller[1][eq][k], other[1][eq][k]), (eq, k)
        for key, v in caller[2].items():
            assert np.array_equal(v, other[2][key], equal_nan=True), key
        assert np.array_equal(caller[3], other[3])


def test_cell_order_on_the_shock_tube():
    """The 1D shock tube (the fixtures' mesh, numbered along the tube): AUTO
    keeps the caller's order; a forced reverse Cuthill-McKee or Hilbert-curve
    order gives the same two steps bit for bit."""
    d = np.load(os.path.join(GOLDEN, "sod_split_2steps.npz"))
    mesh = np.load(os.path.join(GOLDEN, "flow_sod1000_mesh.npz"))
    cfg = dict(FLOW_META["sod64"])
    cfg.update(n_cells=SOD_META["n_cells"], dt=SOD_META["dt"], inner=SOD_META["inner"], solver_type=0,
               solver_mode=1)
    res = {}
    for order in (flow.CELL_ORDER_AUTO, flow.CELL_ORDER_LOCALITY, flow.CELL_ORDER_CURVE):
        prm = params_for(cfg)
        prm.cell_order = order
        with flow.Flow(flow.mesh_from_npz(mesh), d["row_ptr"], d["col"], prm) as fl:
            co = fl.cell_order()
            fl.initialize()
            fl.run(2, SOD_META["inner"])
            res[order] = (co, fl.get_state(), fl.solve_history())
    assert np.array_equal(res[flow.CELL_ORDER_AUTO][0], np.arange(SOD_META["n_cells"]))
    for order in (flow.CELL_ORDER_LOCALITY, flow.CELL_ORDER_CURVE):
        for key, v in res[flow.CELL_ORDER_AUTO][1].items():
            assert np.array_equal(v, res[order][1][key], equal_nan=True), (order, key)
        assert np.array_equal(res[flow.CELL_ORDER_AUTO][2], res[order][2], equal_nan=True), order
