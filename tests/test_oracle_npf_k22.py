"""The oracle's K22 (NPF row-direction conductivity): hy_eff
(gwf-npf.f90:2280-2355) without rotation angles gives K11 on column-direction
faces (j +- 1) and K22 on row-direction faces (i +- 1), the rotation being the
identity (sin 0 = 0, cos 0 = 1 exactly).  Pinned by a closed form and by
transposition: the row faces of a grid equal the column faces of its
transpose with K11 and K22 swapped, so K22 must reach exactly the row faces
with the formula K11 uses on the column faces."""
import numpy as np
import pytest

from oracle import oracle as O

L = O.lib()


def _faces(nlay, nrow, ncol, delr, delc, top2d, botm, k11, k22, k33):
    d = O.dis_connections(nlay, nrow, ncol, delr, delc, top2d, botm)
    cs = np.zeros(d.hwva.size)
    L.orc_calc_condsat(d.nodes, d.ia, d.ja, d.jas, d.ihc, d.cl1, d.cl2, d.hwva, d.top, d.bot,
                       np.ascontiguousarray(k11), np.ascontiguousarray(k22), ncol,
                       np.ascontiguousarray(k33), 0, np.ones(d.nodes),
                       np.ascontiguousarray(d.top), cs)
    out = {}
    for n in range(d.nodes):
        for ii in range(d.ia[n] + 1, d.ia[n + 1]):
            m = int(d.ja[ii])
            if m > n:
                out[(n, m)] = cs[d.jas[ii]]
    return out


def test_k22_closed_form():
    nrow, ncol = 4, 5
    top2d = np.zeros((nrow, ncol))
    botm = np.full((1, nrow, ncol), -5.0)
    n = nrow * ncol
    f = _faces(1, nrow, ncol, np.full(ncol, 10.0), np.full(nrow, 20.0), top2d, botm,
               np.full(n, 3.0), np.full(n, 0.5), np.full(n, 0.1))
    for (a, b), c in f.items():
        if b - a == 1:
            assert np.isclose(c, 3.0 * 5.0 * 20.0 / 10.0, rtol=1e-15)    # column face
        else:
            assert np.isclose(c, 0.5 * 5.0 * 10.0 / 20.0, rtol=1e-15)    # row face


def test_k22_rows_equal_columns_of_the_transpose():
    rng = np.random.default_rng(5)
    nlay, nrow, ncol = 2, 5, 7
    delr, delc = rng.uniform(5, 20, ncol), rng.uniform(5, 20, nrow)
    top2d = rng.uniform(0, 2, (nrow, ncol))
    botm = np.stack([top2d - 4.0 - rng.uniform(0, 1, (nrow, ncol)),
                     top2d - 9.0 - rng.uniform(0, 1, (nrow, ncol))])
    shape = (nlay, nrow, ncol)
    k11 = rng.lognormal(0, 1, shape)
    k22 = rng.lognormal(0, 1, shape)
    k33 = rng.lognormal(-2, 1, shape)
    f = _faces(nlay, nrow, ncol, delr, delc, top2d, botm, k11.ravel(), k22.ravel(), k33.ravel())
    t = lambda a: np.ascontiguousarray(np.swapaxes(a, -1, -2))
    ft = _faces(nlay, ncol, nrow, delc, delr, t(top2d), t(botm), t(k22).ravel(),
                t(k11).ravel(), t(k33).ravel())
    node = lambda k, i, j, nr, nc: (k * nr + i) * nc + j
    nrows = 0
    for k in range(nlay):
        for i in range(nrow):
            for j in range(ncol):
                a = node(k, i, j, nrow, ncol)
                at = node(k, j, i, ncol, nrow)
                if i + 1 < nrow:     # row face here, column face in the transpose
                    assert f[(a, node(k, i + 1, j, nrow, ncol))] == ft[(at, node(k, j, i + 1, ncol, nrow))]
                    nrows += 1
                if j + 1 < ncol:
                    assert f[(a, node(k, i, j + 1, nrow, ncol))] == ft[(at, node(k, j + 1, i, ncol, nrow))]
    assert nrows > 0
    # and K22 matters: with K22 = K11 the row faces change
    f1 = _faces(nlay, nrow, ncol, delr, delc, top2d, botm, k11.ravel(), k11.ravel(), k33.ravel())
    assert f1[(0, ncol)] != f[(0, ncol)] and f1[(0, 1)] == f[(0, 1)]


R = O.ref_gwf_lib()


@pytest.mark.skipif(R is None, reason="oracle/_ref not built")
def test_hyeff_matches_reference():
    """hyeff (HGeoUtil.f90:29-108) bit for bit against the reference, for
    the DIS face normals and random directions, both averagings, angles
    converted with the reference's DPIO180."""
    rng = np.random.default_rng(2)
    axes = [(1.0, 0.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, -1.0, 0.0),
            (0.0, 0.0, 1.0), (0.0, 0.0, -1.0)]
    for _ in range(600):
        k = rng.lognormal(0, 1.5, 3)
        deg = rng.uniform(-180, 180, 3) * (rng.random(3) < 0.8)
        ang = [R.ref_dpio180(a) for a in deg]
        assert all(a == d * O.DPIO180 for a, d in zip(ang, deg))
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        for vg in axes + [tuple(v)]:
            for iavg in (0, 1):
                a = L.orc_hyeff(*k, *ang, *vg, iavg)
                b = R.ref_hyeff(*k, *ang, *vg, iavg)
                assert np.float64(a).view(np.int64) == np.float64(b).view(np.int64)
    # no rotation: the axes give K11, K22, K33 exactly
    assert L.orc_hyeff(2.0, 3.0, 5.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0) == 3.0


def _thick_line(thickstrt, newton=False):
    from modflow6_amd.configs import GwfCase, ImsOptions, Package
    ncol = 6
    return GwfCase(
        name="line_thickstrt", nlay=1, nrow=1, ncol=ncol, delr=np.full(ncol, 10.0),
        delc=np.full(1, 20.0), top2d=np.zeros((1, ncol)), botm=np.full((1, 1, ncol), -10.0),
        k11=np.full(ncol, 3.0), k33=np.full(ncol, 3.0),
        icelltype=np.full(ncol, -1, dtype=np.int32), strt=np.full(ncol, -4.0),
        packages=[Package("CHD", np.array([0, ncol - 1]), np.array([-4.0, -5.0]))],
        ims=ImsOptions(), periods=[(1.0, 1, True)], newton=newton, thickstrt=thickstrt)


def test_thickstrt_closed_form():
    """THICKSTRT (gwf-npf.f90:1836-1886): ICELLTYPE -1 cells become confined
    with the saturated thickness of their starting head, here 6 of 10 m, so
    each face conducts K w b / L = 3 * 20 * 6 / 10; without the option they
    are convertible and calc_condsat uses the full 10 m."""
    s = O.GwfSimulation(_thick_line(True))
    assert np.all(s.icelltype == 0)
    h = s.dis.ihc[s.dis.jas] != 0
    assert np.allclose(s.condsat, 3.0 * 20.0 * 6.0 / 10.0, rtol=1e-15)
    s2 = O.GwfSimulation(_thick_line(False))
    assert np.all(s2.icelltype == 1)
    assert np.allclose(s2.condsat, 3.0 * 20.0 * 10.0 / 10.0, rtol=1e-15)
    assert h.any()
