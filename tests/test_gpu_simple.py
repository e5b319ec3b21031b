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
