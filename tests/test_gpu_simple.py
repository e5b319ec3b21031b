"""The SIMPLE pressure-correction step on the device against the oracle's
restatement of models/single_phase.py:274-365 (scipy spsolve for p')."""

import numpy as np
import pytest

from oracle import fvm as ofvm

pytestmark = pytest.mark.gpu


def _setup(kind, seed):
    from kcfd_amd import workloads
    mesh = workloads.hex_mesh(10) if kind == "hex" else workloads.quad_mesh(24)
    rng = np.random.default_rng(seed)
    n, nd = mesh.n_cells, mesh.ndim
    aP = [np.exp(rng.normal(0.0, 0.5, n)) for _ in range(nd)]
    U = rng.normal(0.0, 1.0, (n, nd))
    p = rng.normal(0.0, 1.0, n)
    mf = rng.normal(0.0, 0.1, mesh.n_faces)
    return mesh, aP, U, p, mf


@pytest.mark.parametrize("kind,dirichlet", [("hex", ["outlet"]), ("quad", ["right"]), ("quad", [])])
def test_pressure_correction_step(ctx, kind, dirichlet):
    from kcfd_amd import device, simple
    from kcfd_amd._lib import DevBuf
    mesh, aP, U, p, mf, = _setup(kind, 11)
    dirichlet = [d for d in dirichlet if d in mesh.boundary_patches]
    pp_o, U_o, p_o, A_o, b_o, dP_o = ofvm.simple_pressure_correction_step(mesh, aP, U, p, mf, 1.2, 0.3,
                                                                         dirichlet)
    dm = device.DeviceMesh(mesh, ctx)
    aPd = [DevBuf.from_array(ctx, a) for a in aP]
    Ud = DevBuf.from_array(ctx, U.reshape(-1))
    pdv = DevBuf.from_array(ctx, p)
    mfd = DevBuf.from_array(ctx, mf)
    st = simple.pressure_correction_step(dm, aPd, Ud, pdv, mfd, 1.2, 0.3, dirichlet, precond="ilu0",
                                         order="color", rtol=1e-13)
    assert np.array_equal(st.dP.download(), dP_o)                       # same operations
    Ag = st.matrix.to_scipy()
    assert np.array_equal(Ag.indptr, A_o.indptr) and np.array_equal(Ag.indices, A_o.indices)
    assert np.max(np.abs(Ag.data - A_o.data)) <= 1e-14 * np.max(np.abs(A_o.data))
    assert np.max(np.abs(st.matrix.get_rhs() - b_o)) <= 1e-15 * max(1.0, np.max(np.abs(b_o)))
    pp = st.pp.download()
    e_pp = np.linalg.norm(pp - pp_o) / np.linalg.norm(pp_o)
    e_U = np.max(np.abs(Ud.download().reshape(U.shape) - U_o)) / np.max(np.abs(U_o))
    e_p = np.max(np.abs(pdv.download() - p_o)) / np.max(np.abs(p_o))
    print(f"{kind} {dirichlet}: solve {st.solve['iters']} it, info {st.solve['info']}; "
          f"p' rel {e_pp:.2e}, U {e_U:.2e}, p {e_p:.2e}")
    assert st.solve["info"] == 0
    assert e_pp < 1e-10 and e_U < 1e-10 and e_p < 1e-10
    ratio_o = np.linalg.norm(pp_o) / max(np.linalg.norm(p_o) + 1e-10, 1e-10)   # single_phase.py:376
    assert abs(st.ratio - ratio_o) <= 1e-9 * ratio_o


def test_correction_matches_oracle(ctx):
    """Given p', the fused correction equals the oracle's Green-Gauss +
    update (same face-order sums; the interpolation weights come from the
    device mesh's precomputed geometry, so the bar is the Green-Gauss
    kernel's, 1e-13)."""
    from kcfd_amd import device
    from kcfd_amd._lib import DevBuf, check, lib
    mesh, aP, U, p, mf = _setup("hex", 12)
    rng = np.random.default_rng(3)
    pp = rng.normal(0, 1, mesh.n_cells)
    dP = rng.uniform(0.5, 2.0, mesh.n_cells)
    dirichlet = ["outlet"]
    bv = {nm: (np.zeros(len(f)) if nm in dirichlet else pp[mesh.owner[np.asarray(f)]])
          for nm, f in mesh.boundary_patches.items()}
    g = ofvm.green_gauss_gradient(mesh, pp, bv)
    U_o = U - dP[:, None] * g
    p_o = p + 0.3 * pp
    dm = device.DeviceMesh(mesh, ctx)
    Ud, pdv = DevBuf.from_array(ctx, U.reshape(-1)), DevBuf.from_array(ctx, p)
    pd = np.array([1 if nm in dirichlet else 0 for nm in dm.patch_names], np.int32)
    ppd, dPd = DevBuf.from_array(ctx, pp), DevBuf.from_array(ctx, dP)   # held: freed buffers get reused
    check(lib().kcfd_simple_correct(dm.h, ppd.p, dPd.p, pd.ctypes.data, 0.3, Ud.p, pdv.p))
    assert np.max(np.abs(Ud.download().reshape(U.shape) - U_o)) <= 1e-13 * np.max(np.abs(U_o))
    assert np.array_equal(pdv.download(), p_o)


def _tf_setup(kind, seed):
    from kcfd_amd import workloads
    mesh = {"hex": lambda: workloads.hex_mesh(10), "quad": lambda: workloads.quad_mesh(24),
            "hybrid": lambda: workloads.hybrid_mesh(8, 3, 3)}[kind]()
    rng = np.random.default_rng(seed)
    n, nd, nf = mesh.n_cells, mesh.ndim, mesh.n_faces
    al = rng.uniform(0.05, 0.95, n)
    f = dict(alpha_l=al, alpha_g=1.0 - al, aP_l=np.exp(rng.normal(0, 0.5, n)),
             aP_g=0.1 * np.exp(rng.normal(0, 0.5, n)), mf_l=rng.normal(0, 0.1, nf),
             mf_g=rng.normal(0, 0.01, nf))
    st = dict(U_l=rng.normal(0, 1, (n, nd)), U_g=rng.normal(0, 1, (n, nd)), p=rng.normal(0, 1, n))
    return mesh, f, st


@pytest.mark.parametrize("kind", ["hex", "quad", "hybrid"])
def test_two_fluid_pressure_correction_step(ctx, kind):
    """models/two_fluid.py:641-725 on the device against the oracle (spsolve
    for p'): matrix within 1e-14, right-hand side within 1e-15, p', both
    velocities and p within 1e-9, and the reference's return ratio."""
    from kcfd_amd import device, simple
    from kcfd_amd._lib import DevBuf
    mesh, f, st = _tf_setup(kind, 21)
    rho_l, rho_g, alpha_p = 998.2, 1.225, 0.3
    pp_o, Ul_o, Ug_o, p_o, A_o, b_o, g_o, ratio_o = ofvm.two_fluid_pressure_correction_step(
        mesh, f["alpha_l"], f["alpha_g"], f["aP_l"], f["aP_g"], rho_l, rho_g, f["mf_l"], f["mf_g"],
        st["U_l"], st["U_g"], st["p"], alpha_p)
    dm = device.DeviceMesh(mesh, ctx)
    d = {k: DevBuf.from_array(ctx, v) for k, v in f.items()}
    Uld, Ugd = DevBuf.from_array(ctx, st["U_l"].reshape(-1)), DevBuf.from_array(ctx, st["U_g"].reshape(-1))
    pdv = DevBuf.from_array(ctx, st["p"])
    res = simple.two_fluid_pressure_correction_step(dm, d["alpha_l"], d["alpha_g"], d["aP_l"], d["aP_g"],
                                                    rho_l, rho_g, d["mf_l"], d["mf_g"], Uld, Ugd, pdv,
                                                    alpha_p, precond="ilu0", order="color", rtol=1e-13)
    Ag = res.matrix.to_scipy()
    assert np.array_equal(Ag.indptr, A_o.indptr) and np.array_equal(Ag.indices, A_o.indices)
    ea = np.max(np.abs(Ag.data - A_o.data) / np.abs(A_o.data))
    eb = np.max(np.abs(res.matrix.get_rhs() - b_o)) / max(1.0, np.max(np.abs(b_o)))
    pp = res.pp.download()
    e_pp = np.linalg.norm(pp - pp_o) / np.linalg.norm(pp_o)
    sh = st["U_l"].shape
    e_ul = np.max(np.abs(Uld.download().reshape(sh) - Ul_o)) / np.max(np.abs(Ul_o))
    e_ug = np.max(np.abs(Ugd.download().reshape(sh) - Ug_o)) / np.max(np.abs(Ug_o))
    e_p = np.max(np.abs(pdv.download() - p_o)) / np.max(np.abs(p_o))
    print(f"two-fluid {kind}: A {ea:.1e} b {eb:.1e}; solve {res.solve['iters']} it info "
          f"{res.solve['info']}; p' {e_pp:.1e} U_l {e_ul:.1e} U_g {e_ug:.1e} p {e_p:.1e}; ratio "
          f"{res.ratio:.6e} vs {ratio_o:.6e}")
    assert ea <= 1e-14 and eb <= 1e-15
    assert res.solve["info"] == 0
    assert e_pp < 1e-9 and e_ul < 1e-9 and e_ug < 1e-9 and e_p < 1e-9
    assert abs(res.ratio - ratio_o) <= 1e-9 * ratio_o


def test_two_fluid_correction_matches_oracle(ctx):
    """Given p', kcfd_twofluid_correct equals the oracle's face-difference
    correction of both phases and the pressure update (same face-order
    sums)."""
    from kcfd_amd import device
    from kcfd_amd._lib import DevBuf, check, lib
    mesh, f, st = _tf_setup("hybrid", 22)
    pp = np.random.default_rng(4).normal(0, 1, mesh.n_cells)
    Ul_o, Ug_o, p_o, _ = ofvm.two_fluid_correction(mesh, pp, f["alpha_l"], f["alpha_g"], f["aP_l"], f["aP_g"],
                                                   st["U_l"], st["U_g"], st["p"], 0.4)
    dm = device.DeviceMesh(mesh, ctx)
    d = {k: DevBuf.from_array(ctx, v) for k, v in f.items()}
    ppd = DevBuf.from_array(ctx, pp)
    Uld, Ugd = DevBuf.from_array(ctx, st["U_l"].reshape(-1)), DevBuf.from_array(ctx, st["U_g"].reshape(-1))
    pdv = DevBuf.from_array(ctx, st["p"])
    check(lib().kcfd_twofluid_correct(dm.h, ppd.p, d["alpha_l"].p, d["alpha_g"].p, d["aP_l"].p, d["aP_g"].p,
                                      0.4, Uld.p, Ugd.p, pdv.p))
    sh = st["U_l"].shape
    for got, want in [(Uld.download().reshape(sh), Ul_o), (Ugd.download().reshape(sh), Ug_o)]:
        assert np.max(np.abs(got - want)) <= 1e-14 * np.max(np.abs(want))
    assert np.array_equal(pdv.download(), p_o)
