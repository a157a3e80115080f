"""The package's own HDF5 reader on two of the reference's mesh files (CPU).

tests/golden/pLaplace_level1.h5 and GL_level2.h5 are copies of the reference's
data/meshes files (data, written by h5py).  The reader is pinned by what the parsed arrays
must satisfy if every dataset was decoded right (dtype, shape, order, layout):

* vol is half the absolute determinant of each triangle's edge vectors, from nodes and elems;
* dvx, dvy are the gradients of the P1 basis: sum over the vertices is 0 and
  sum_a x_a dvx_a = 1, sum_a y_a dvx_a = 0 (and the same for y);
* the adjacency is the free-DOF pattern of the elements (symmetric, diagonal included);
* free DOFs are the nodes off the mesh boundary (edges of one triangle), and the p-Laplace u_0 is
  zero on that boundary;
* the GL quadrature table: barycentric points (columns sum to 1), weights summing to 1, eps 0.01;
* the arrays equal the committed npz fixtures written from the same files.
"""

from pathlib import Path

import numpy as np
import pytest

from fenics_nonlinear_energies_amd.hdf5 import load_problem_hdf5
from tests import fixtures

GOLD = Path(__file__).resolve().parent / "golden"


def _geometry_checks(p):
    nodes = np.asarray(p["nodes"], dtype=np.float64)
    elems = np.asarray(p["elems"], dtype=np.int64)
    xy = nodes[elems]  # (N, 3, 2)
    e1, e2 = xy[:, 1] - xy[:, 0], xy[:, 2] - xy[:, 0]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    assert np.allclose(p["vol"], area, rtol=1e-12, atol=0)
    dvx, dvy = np.asarray(p["dvx"]), np.asarray(p["dvy"])
    assert np.allclose(dvx.sum(1), 0, atol=1e-10) and np.allclose(dvy.sum(1), 0, atol=1e-10)
    x, y = xy[..., 0], xy[..., 1]
    assert np.allclose((x * dvx).sum(1), 1, atol=1e-10) and np.allclose((y * dvx).sum(1), 0, atol=1e-10)
    assert np.allclose((x * dvy).sum(1), 0, atol=1e-10) and np.allclose((y * dvy).sum(1), 1, atol=1e-10)


def _adjacency_checks(p, adj):
    fd = np.asarray(p["freedofs"], dtype=np.int64).ravel()
    n_total = np.asarray(p["u_0"]).size
    f_of = np.full(n_total, -1)
    f_of[fd] = np.arange(fd.size)
    ef = f_of[np.asarray(p["elems"], dtype=np.int64)]
    keys = {(int(a), int(b)) for row in ef for a in row for b in row if a >= 0 and b >= 0}
    a = adj.tocoo()
    got = {(int(r), int(c)) for r, c in zip(a.row, a.col)}
    assert got == keys


@pytest.mark.parametrize("name,npz", [("pLaplace_level1.h5", "plaplace_l1"), ("GL_level2.h5", "gl_l2")])
def test_reference_mesh_files_decode(name, npz):
    p, adj = load_problem_hdf5(str(GOLD / name))
    _geometry_checks(p)
    _adjacency_checks(p, adj)
    # free DOFs are the nodes off the mesh boundary (edges of exactly one triangle)
    elems = np.asarray(p["elems"], dtype=np.int64)
    edges = np.sort(np.concatenate([elems[:, [0, 1]], elems[:, [1, 2]], elems[:, [2, 0]]]), axis=1)
    uniq, count = np.unique(edges, axis=0, return_counts=True)
    boundary = np.zeros(np.asarray(p["nodes"]).shape[0], dtype=bool)
    boundary[uniq[count == 1].ravel()] = True
    fd = np.asarray(p["freedofs"], dtype=np.int64).ravel()
    assert set(fd.tolist()) == set(np.nonzero(~boundary)[0].tolist())
    ref = fixtures.plaplace(1)[0] if npz == "plaplace_l1" else fixtures.gl(2)[0]
    for k, v in ref.items():
        if k in p:
            assert np.array_equal(np.asarray(p[k]).ravel(), np.asarray(v).ravel()), k
    if npz == "gl_l2":
        ip, w = np.asarray(p["ip"]), np.asarray(p["w"])
        assert np.allclose(ip.sum(0), 1.0, atol=1e-12) and np.isclose(w.sum(), 1.0) and float(p["eps"]) == 0.01
    else:
        assert not np.asarray(p["u_0"])[boundary].any()


def test_not_an_hdf5_file_is_rejected(tmp_path):
    from fenics_nonlinear_energies_amd.hdf5 import HDF5FormatError

    bad = tmp_path / "bad.h5"
    bad.write_bytes(b"not an hdf5 file" * 8)
    with pytest.raises(HDF5FormatError):
        load_problem_hdf5(str(bad))
