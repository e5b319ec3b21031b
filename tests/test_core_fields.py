"""The drop-in field containers (core/fields.py:12-155) on the host: the
VectorField component views x / y / z (z reads zeros and ignores writes on
a 2-D mesh), set_uniform on cells and boundary faces, a short default
zero-padded to the mesh's dimension, and copies that share nothing."""

import numpy as np


def _meshes():
    from kcfd_amd import workloads
    return workloads.quad_mesh(6), workloads.hex_mesh(3)


def test_vector_components_2d_and_3d():
    from kcfd_amd import core
    quad, hexm = _meshes()
    v2 = core.VectorField(quad, "U", default=[1.0, 2.0])
    assert v2.values.shape == (quad.n_cells, 2)
    assert np.all(v2.x == 1.0) and np.all(v2.y == 2.0)
    assert np.all(v2.z == 0.0) and v2.z.shape == (quad.n_cells,)
    v2.z = 5.0                                  # ignored in 2-D
    assert v2.values.shape[1] == 2 and np.all(v2.z == 0.0)
    v2.x = np.arange(quad.n_cells)
    assert np.array_equal(v2.values[:, 0], np.arange(quad.n_cells))
    v2.x[0] = -1.0                              # x is a view of values
    assert v2.values[0, 0] == -1.0

    v3 = core.VectorField(hexm, "U", default=[1.0])   # short default, padded
    assert v3.values.shape == (hexm.n_cells, 3)
    assert np.all(v3.values == np.array([1.0, 0.0, 0.0]))
    for bv in v3.boundary_values.values():
        assert bv.shape[1] == 3 and np.all(bv == np.array([1.0, 0.0, 0.0]))
    v3.z = 7.0
    assert np.all(v3.values[:, 2] == 7.0)


def test_vector_set_uniform_and_copy():
    from kcfd_amd import core
    _, hexm = _meshes()
    v = core.VectorField(hexm, "U")
    v.set_uniform([0.5, -1.0, 2.0])
    assert np.all(v.values == np.array([0.5, -1.0, 2.0]))
    for bv in v.boundary_values.values():
        assert np.all(bv == np.array([0.5, -1.0, 2.0]))
    v.store_old()
    c = v.copy()
    c.x = 9.0
    c.boundary_values[next(iter(c.boundary_values))][:] = 3.0
    c.old_values[:] = 4.0
    assert np.all(v.x == 0.5) and np.all(v.old_values[:, 0] == 0.5)
    assert all(np.all(bv[:, 0] == 0.5) for bv in v.boundary_values.values())
    assert np.allclose(v.magnitude(), np.sqrt(0.25 + 1.0 + 4.0))


def test_scalar_set_uniform():
    from kcfd_amd import core
    quad, _ = _meshes()
    s = core.ScalarField(quad, "p", default=1.5)
    assert np.all(s.values == 1.5)
    s.set_uniform(-2.0)
    assert np.all(s.values == -2.0) and all(np.all(b == -2.0) for b in s.boundary_values.values())
