"""MatrixMarket files (LinearAlgebraInterface::writeToFile,
linearAlgebraInterface_util.hpp:316-331).  The files written by the engine
are read by an independent reader (scipy.io.mmread) and by the engine's
own reader, bit for bit; files written by scipy (general and symmetric) are
placed into the CRS pattern; an entry outside the pattern is refused."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from conftest import to_scipy
from mrhyde_amd.abi import Problem

pytestmark = pytest.mark.gpu
PI = np.pi


@pytest.fixture(scope="module")
def ctx():
    from mrhyde_amd import Mhx
    from oracle import Oracle

    p = Problem(dim=2, order=2, quadrature=4, n=(5, 4), diff_u2=1.0, modes=[(3.0, (PI, PI))], perturb=0.2,
                perturb_seed=1, dbc_sides=1 | 2)
    g, o = Mhx(p), Oracle(p)
    u = o.setDirichlet(np.random.default_rng(3).standard_normal(o.n_dof))
    r, j = np.zeros(g.n_dof), np.zeros(g.nnz)
    g.assemble(u, r, j)
    return g, o, u, r, j


def test_write_then_read_bit_exact(ctx, tmp_path):
    g, o, u, r, j = ctx
    jf, rf, sf = (str(tmp_path / n) for n in ("J.mm", "r.mm", "x.mm"))
    g.writeToFile(jac=j, res=r, sol=u, jac_filename=jf, res_filename=rf, sol_filename=sf)
    rp, ci = g.crs_pattern()
    A = scipy.io.mmread(jf).tocsr()
    A.sort_indices()
    ref = to_scipy(rp, ci, j)
    np.testing.assert_array_equal(A.indptr, ref.indptr)
    np.testing.assert_array_equal(A.indices, ref.indices)
    np.testing.assert_array_equal(A.data, ref.data)
    np.testing.assert_array_equal(scipy.io.mmread(rf).ravel(), r)
    np.testing.assert_array_equal(scipy.io.mmread(sf).ravel(), u)
    np.testing.assert_array_equal(g.readMatrixMarket(jf), j)
    np.testing.assert_array_equal(g.readVectorMarket(rf), r)


def test_read_scipy_files(ctx, tmp_path):
    g, o, u, r, j = ctx
    rp, ci = g.crs_pattern()
    A = to_scipy(rp, ci, j)
    f = str(tmp_path / "scipy_general.mtx")
    scipy.io.mmwrite(f, A.tocoo(), precision=17)
    np.testing.assert_array_equal(g.readMatrixMarket(f), j)
    # a symmetric matrix on the same pattern, written in symmetric form
    S = (A + A.T) * 0.5
    S = sp.csr_matrix((np.asarray(S[S.nonzero()]).ravel(), S.nonzero()), shape=A.shape)
    fs = str(tmp_path / "scipy_sym.mtx")
    scipy.io.mmwrite(fs, S.tocoo(), symmetry="symmetric", precision=17)
    got = to_scipy(rp, ci, g.readMatrixMarket(fs))
    assert abs(got - S).max() <= 1e-15 * abs(S).max()
    fv = str(tmp_path / "vec.mtx")
    scipy.io.mmwrite(fv, u.reshape(-1, 1), precision=17)
    np.testing.assert_array_equal(g.readVectorMarket(fv), u)


def test_entry_outside_pattern_is_refused(ctx, tmp_path):
    from mrhyde_amd import MhxError

    g, o, u, r, j = ctx
    n = g.n_dof
    f = str(tmp_path / "bad.mm")
    with open(f, "w") as fh:
        fh.write("%%MatrixMarket matrix coordinate real general\n")
        fh.write(f"{n} {n} 1\n1 {n} 1.0\n")  # the first and last DOF share no element
    with pytest.raises(MhxError):
        g.readMatrixMarket(f)
