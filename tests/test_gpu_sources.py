"""GPU parity of the lattice sources and gausslet set-up (rpx_lattice_*,
rpx_config_parabasal_rays, rpx_project_to_plane) against the oracle's numpy /
C restatements of sources.py:1035-1107, gausslet_sources.py:386-457 and
516-566, and ctracer.pyx:1391-1481.

Bar: counts, neighbour tables and every geometric field bit-exact.  The
amplitudes carry the Gaussian weight exp(...) -- the device's exponential
against numpy's, each within an ulp or two of the true value -- so E1/E2 are
compared to 4e-16 relative; for the collimated source the power scaling
is a sum over all gausslets whose order differs from numpy's pairwise sum,
so there the bound is 1e-14 relative.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import sources as osrc
from raypier_optics_amd import workloads
from raypier_optics_amd._lib import RpxError
from raypier_optics_amd.ctracer import GaussletCollection
from raypier_optics_amd.sources import HexagonalRayFieldSource, CollimatedGaussletSource, GaussianPointSource

pytestmark = pytest.mark.gpu

AMP = ("E1_amp", "E2_amp")


def _base(a):
    return a["base_ray"] if "base_ray" in a.dtype.names else a


def assert_rays(got, ref, amp_rtol, what):
    assert len(got) == len(ref), (what, len(got), len(ref))
    g, r = _base(got), _base(ref)
    for f in r.dtype.names:
        if f in AMP:
            scale = np.maximum(np.abs(r[f]), 1e-300)
            err = np.abs(g[f] - r[f]) / scale
            assert err.max() <= amp_rtol, (what, f, err.max())
        else:
            assert g[f].tobytes() == r[f].tobytes(), (what, f)
    if "para_rays" in ref.dtype.names:
        assert got["para_rays"].tobytes() == ref["para_rays"].tobytes(), (what, "para_rays")


def test_hex_ray_field_known_answer(gpu):
    """test_fields.py:35-48: radius 10 at spacing 4 gives 19 rays, 7 of them
    with all six neighbours."""
    src = HexagonalRayFieldSource(origin=(0, 0, -50), direction=(0, 0, 1), radius=10.0, resolution=2.5)
    rc = src.input_rays
    assert rc.n_rays == 19
    nb = rc.neighbours
    assert nb.shape == (19, 6)
    assert int((nb >= 0).all(axis=1).sum()) == 7
    # every neighbour is one lattice spacing away, and the relation is symmetric
    o = rc.origin
    for i in range(19):
        for k in range(6):
            j = nb[i, k]
            if j >= 0:
                assert abs(np.linalg.norm(o[j] - o[i]) - 4.0) < 1e-12
                assert nb[j, (k + 3) % 6] == i


@pytest.mark.parametrize("cfg", [
    dict(origin=(0, 0, -50), direction=(0, 0, 1), radius=10.0, resolution=10.0),
    dict(origin=(1.5, -2.0, 3.0), direction=(0.1, 0.2, 1.0), radius=6.0, resolution=37.3, gauss_width=3.0,
         E_vector=(0.0, 1.0, 0.0), E1_amp=0.6 + 0.3j, E2_amp=-0.2 + 0.7j),
    dict(origin=(0, 0, 0), direction=(1.0, 0.3, 0.1), radius=25.0, resolution=300.0),
])
def test_hex_ray_field_parity(gpu, cfg):
    src = HexagonalRayFieldSource(**cfg)
    rc = src.input_rays
    ref, ref_nb = osrc.hexagonal_ray_field(src.origin, src.direction, src.radius, src.resolution, src.gauss_width,
                                           src.E_vector, src.E1_amp, src.E2_amp)
    assert_rays(rc.copy_as_array(), ref, 4e-16, "hex %s" % cfg)
    assert np.array_equal(rc.neighbours, ref_nb)
    assert list(rc.wavelengths) == [src.wavelength]


@pytest.mark.parametrize("cfg", [
    dict(origin=(0, 0, -80), radius=4.0, resolution=40.0, beam_waist=2.0, wavelength=1.064),
    dict(origin=(0.5, 0.2, -10), direction=(0.05, -0.1, 1.0), radius=3.0, resolution=123.4, beam_waist=1.5,
         blending=2.0, wavelength=0.6328, beam_power=3.5, E1_amp=0.8 + 0.1j, E2_amp=0.3j),
])
def test_collimated_gausslets_parity(gpu, cfg):
    src = CollimatedGaussletSource(**cfg)
    gc = src.input_rays
    ref = osrc.collimated_gausslets(src.origin, src.direction, src.radius, src.resolution, src.beam_waist,
                                    src.blending, src.wavelength, src.beam_power, src.E_vector, src.E1_amp,
                                    src.E2_amp)
    got = gc.copy_as_array()
    assert_rays(got, ref, 1e-14, "collimated %s" % cfg)
    b = got["base_ray"]
    E1, E2 = complex(src.E1_amp), complex(src.E2_amp)
    P = (E1 * E1.conjugate()).real + (E2 * E1.conjugate()).real
    g2 = np.abs(b["E1_amp"]) ** 2 / abs(E1) ** 2
    assert abs((P * g2).sum() * 1.0 - src.beam_power) <= 1e-12 * src.beam_power


@pytest.mark.parametrize("cfg", [
    dict(resolution=10.0, numerical_aperture=0.2, working_dist=100.0, wavelength=0.78, beam_waist=1.0),
    dict(origin=(1.0, 2.0, 3.0), direction=(0.0, 0.3, 1.0), resolution=25.5, numerical_aperture=0.12,
         working_dist=-20.0, wavelength=1.064, beam_waist=5.0, blending=1.7, E_vector=(0.0, 1.0, 0.0),
         E1_amp=0.5 - 0.5j, E2_amp=0.25 + 0j),
])
def test_point_source_parity(gpu, cfg):
    src = GaussianPointSource(**cfg)
    gc = src.input_rays
    ref = osrc.gaussian_point_source(src.origin, src.direction, src.resolution, src.working_dist,
                                     src.numerical_aperture, src.blending, src.beam_waist, src.wavelength,
                                     src.E_vector, src.E1_amp, src.E2_amp)
    assert_rays(gc.copy_as_array(), ref, 4e-16, "point %s" % cfg)


def test_point_source_known_count(gpu):
    """make_hexagonal_grid at 2.5 spacings: the lattice points closer than
    2.5 spacings to the centre -- distances 0, 1, sqrt(3) and 2 -- are 19."""
    gc = GaussianPointSource(resolution=2.5).input_rays
    assert gc.n_rays == 19
    d = gc.copy_as_array()["base_ray"]["direction"]
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, rtol=1e-15)
    assert np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2).max() < 0.2


def _random_gausslets(n, seed):
    rng = np.random.default_rng(seed)
    r = workloads.base_rays(n)
    d = rng.normal(size=(n, 3))
    d[:, 2] = np.abs(d[:, 2]) + 0.5
    r["direction"] = d / np.linalg.norm(d, axis=1)[:, None]
    r["origin"] = rng.uniform(-5, 5, (n, 3))
    r["normal"] = rng.normal(size=(n, 3))
    r["length"] = rng.uniform(1, 9, n)
    r["wavelength_idx"] = rng.integers(0, 3, n)
    r["refractive_index"] = rng.uniform(1.0, 1.7, n) + 1e-3j
    r["accumulated_path"] = rng.uniform(0, 10, n)
    return GaussletCollection.from_rays(r)


def test_config_parabasal_and_project(gpu):
    wl = np.array([0.5, 0.8, 1.064])
    for radius, wd in ((0.05, 0.0), (0.2, 3.5)):
        g = _random_gausslets(5000, 3)
        before = g.copy_as_array()
        g.config_parabasal_rays(wl, radius, wd)
        ref = oracle.config_parabasal_rays(before, wl, radius, wd)
        assert g.copy_as_array().tobytes() == ref.tobytes()
        o, d = (0.3, -0.2, 12.0), (0.1, 0.05, 1.0)
        g.project_to_plane(o, d)
        assert g.copy_as_array().tobytes() == oracle.project_to_plane(ref, o, d).tobytes()


def test_config_parabasal_bad_wavelength(gpu):
    g = _random_gausslets(100, 4)
    with pytest.raises(RpxError):
        g.config_parabasal_rays(np.array([0.5, 0.8]), 0.05, 0.0)


def test_source_traces(gpu):
    """A device-generated ray field goes straight into the trace, and its
    neighbour table follows the generations."""
    from raypier_optics_amd.tracer import trace_rays
    w = workloads.singlet_c1(n_rays=10)
    src = HexagonalRayFieldSource(origin=(0, 0, -10), direction=(0, 0, 1), radius=8.0, resolution=20.0,
                                  wavelength=0.55)
    rc = src.input_rays
    nb0 = rc.neighbours
    traced, _ = trace_rays(rc, w.face_lists, recursion_limit=4, max_length=100.0)
    ref = nb0
    for g in range(1, len(traced)):
        ref = oracle.eval_neighbours(traced[g].copy_as_array(), ref)
        assert np.array_equal(traced[g].neighbours, ref)
    assert traced[2].n_rays > 0.9 * rc.n_rays


def test_lagrange_invariant(gpu):
    """GaussletCollection.lagrange_invariant (ctracer.pyx:1347-1389) on the
    device against the oracle: symmetric parabasal sets from
    config_parabasal_rays and random asymmetric ones (1e-13 relative; the
    value ends in one square root and a division).  Then through a planar
    n = 1.5 interface traced on the device (exactly 1/1.5 of the input) and
    through a lossless singlet back into air (conserved to 2e-5)."""
    from raypier_optics_amd._scene import compile_scene
    from raypier_optics_amd.tracer import trace_rays
    from tests.test_gausslet_host import asymmetric_gausslets, lagrange_scene
    wl = np.array([0.5, 0.8, 1.064])
    g = _random_gausslets(3000, 9)
    g.config_parabasal_rays(wl, 0.05, 1.0)
    g.wavelengths = wl
    np.testing.assert_allclose(g.lagrange_invariant, oracle.lagrange_invariant(g.copy_as_array(), wl), rtol=1e-13,
                               atol=1e-13)
    a = asymmetric_gausslets(4000, 3)
    ga = GaussletCollection.from_array(a)
    ga.wavelengths = np.array([0.55])
    L0 = ga.lagrange_invariant
    np.testing.assert_allclose(L0, oracle.lagrange_invariant(a, [0.55]), rtol=1e-13)
    flat, lens = lagrange_scene()
    traced, _ = trace_rays(ga, [flat], recursion_limit=2, max_length=100.0)
    assert traced[1].n_rays == len(a)
    np.testing.assert_allclose(traced[1].lagrange_invariant, L0 / 1.5, rtol=1e-6)
    gb = GaussletCollection.from_array(a)
    gb.wavelengths = np.array([0.55])
    traced, _ = trace_rays(gb, [lens], recursion_limit=3, max_length=100.0)
    assert traced[2].n_rays == len(a)
    L2 = traced[2].lagrange_invariant
    np.testing.assert_allclose(L2, L0, rtol=2e-5)
    workloads.prepare_faces([lens], [0.55], 100.0)
    cur = a.copy()
    t = compile_scene([lens])
    for _ in range(2):
        cur, _ = oracle.trace_gausslet(t, cur, 100.0, [0.55])
    np.testing.assert_allclose(L2, oracle.lagrange_invariant(cur, [0.55]), rtol=1e-9)
