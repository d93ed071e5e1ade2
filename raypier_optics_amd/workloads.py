"""Optical systems and seeded ray fields of the benchmark configurations,
built with the drop-in API.

  singlet_c1      N-BK7 biconvex singlet (R=+/-50 mm, CT 5 mm, D 25 mm),
                  uncoated Fresnel interfaces with zero thresholds (reflected
                  children kept), opaque detector at the paraxial focus;
                  collimated hexagonal field, 10,000 rays over 20 mm.
  asphere_doublet_c2
                  even asphere (k=-0.6, A4..A10) in N-BK7 followed by a
                  cemented N-BK7/N-SF5 doublet R=(+61.5, -44.2, -128.2) mm,
                  single-layer MgF2 (n=1.38) quarter-wave AR coating on the
                  outer faces, planar detector; 4,194,304 rays uniform on a
                  6 mm-radius disk, angles U(+/-2 deg), per-ray wavelength
                  U(0.45, 0.65) um.

Sign conventions follow raypier: SphericalFace.curvature = z_vertex -
z_centre (= -R for optical radius R, centre at vertex + R); AsphericFace.
curvature = -R with the polynomial terms in the usual sag convention
(z = c r^2/(1+sqrt(1-(1+k)c^2 r^2)) + A4 r^4 + ...); a material's "outside"
is the side its face normal points to.
"""
import math

import numpy as np

from ._abi import ray_dtype, NO_FACE
from .ctracer import RayCollection, FaceList
from .cfaces import SphericalFace, CircularFace, AsphericFace, ShapedPlanarFace
from .cshapes import CircleShape
from .cmaterials import (BaseDispersionCurve, NondispersiveCurve, FullDielectricDispersiveMaterial,
                         CoatedDispersiveMaterial, OpaqueMaterial)

# Sellmeier formula-2 coefficients [0, B1, C1, B2, C2, B3, C3] (C in um^2)
N_BK7 = (0.0, 1.03961212, 0.00600069867, 0.231792344, 0.0200179144, 1.01046945, 103.560653)
N_SF5 = (0.0, 1.52481889, 0.011254756, 0.187085527, 0.0588995392, 1.42729015, 129.141675)
N_SF11 = (0.0, 1.73759695, 0.013188707, 0.313747346, 0.0623068142, 1.89878101, 155.23629)


def glass(coefs, wl_min=0.3, wl_max=2.5):
    return BaseDispersionCurve(2, np.array(coefs), 0.0, wl_min, wl_max)


def sellmeier_n(coefs, wl):
    wl2 = wl * wl
    n2 = 1.0 + coefs[0]
    for i in range(3):
        n2 += coefs[2 * i + 1] * wl2 / (wl2 - coefs[2 * i + 2])
    return math.sqrt(n2)


# ------------------------------------------------------------- ray fields
def base_rays(n):
    r = np.zeros(n, ray_dtype)
    r["direction"] = (0.0, 0.0, 1.0)
    r["E_vector"] = (1.0, 0.0, 0.0)
    r["refractive_index"] = 1.0
    r["E1_amp"] = 1.0
    r["length"] = np.inf
    r["parent_idx"] = NO_FACE
    r["end_face_idx"] = NO_FACE
    r["ray_ident"] = np.arange(n, dtype=np.uint32)
    return r


def hex_field(n_rays, radius, z0):
    """Collimated hexagonal grid inside a circle, trimmed to the n_rays points
    nearest the axis (ties broken by index)."""
    area = math.pi * radius * radius
    spacing = math.sqrt(2.0 * area / (math.sqrt(3.0) * n_rays))
    m = int(radius / spacing) + 3
    i, j = np.meshgrid(np.arange(-m, m + 1), np.arange(-m, m + 1), indexing="ij")
    x = (i + 0.5 * j) * spacing
    y = j * (math.sqrt(3.0) / 2.0) * spacing
    x, y = x.ravel(), y.ravel()
    r2 = x * x + y * y
    order = np.lexsort((np.arange(len(r2)), r2))[:n_rays]
    r = base_rays(n_rays)
    r["origin"][:, 0] = x[order]
    r["origin"][:, 1] = y[order]
    r["origin"][:, 2] = z0
    return r


def disk_field(n_rays, radius, z0, max_angle_deg, wl_range, seed):
    """Uniform positions on a disk, direction angles U(+/-max_angle), one
    wavelength per ray (wavelength_idx = ray index)."""
    rng = np.random.default_rng(seed)
    rr = radius * np.sqrt(rng.random(n_rays))
    th = 2 * np.pi * rng.random(n_rays)
    ax = np.radians(rng.uniform(-max_angle_deg, max_angle_deg, n_rays))
    ay = np.radians(rng.uniform(-max_angle_deg, max_angle_deg, n_rays))
    d = np.stack([np.tan(ax), np.tan(ay), np.ones(n_rays)], axis=1)
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = base_rays(n_rays)
    r["origin"][:, 0] = rr * np.cos(th)
    r["origin"][:, 1] = rr * np.sin(th)
    r["origin"][:, 2] = z0
    r["direction"] = d
    e = np.cross(d, np.array([0.0, 1.0, 0.0]))   # linear polarisation along ~x
    r["E_vector"] = e / np.linalg.norm(e, axis=1)[:, None]
    r["wavelength_idx"] = np.arange(n_rays, dtype=np.uint32)
    wl = rng.uniform(wl_range[0], wl_range[1], n_rays)
    return r, wl


# ---------------------------------------------------------------- systems
class Workload(object):
    def __init__(self, name, rays, wavelengths, face_lists, recursion_limit, max_length, info, source=None):
        self.name = name
        self.rays = rays                  # host ray_dtype array (None for a device source)
        self.wavelengths = np.asarray(wavelengths, np.float64)
        self.face_lists = face_lists
        self.recursion_limit = recursion_limit
        self.max_length = max_length
        self.info = info
        self.source = source              # callable building the input RayCollection on the device

    @property
    def n_rays(self):
        return len(self.rays) if self.rays is not None else self.info["n_rays"]

    def collection(self):
        if self.source is not None:
            rc = self.source()
        else:
            rc = RayCollection.from_array(self.rays)
        rc.wavelengths = self.wavelengths
        return rc

    def subset(self, idx):
        """The same system with a subset of the input rays (wavelength table
        reindexed when rays carry their own wavelength)."""
        rays = self.rays[idx].copy()
        wl = self.wavelengths
        if len(wl) == len(self.rays) and len(wl) > 1:
            wl = wl[rays["wavelength_idx"]]
            rays["wavelength_idx"] = np.arange(len(rays), dtype=np.uint32)
        return Workload(self.name, rays, wl, self.face_lists, self.recursion_limit, self.max_length, self.info)


def prepare_faces(face_lists, wavelengths, max_length):
    """Global face indices, counts and material wavelengths exactly as
    trace_rays sets them (core/tracer.py:28-34); returns all_faces."""
    all_faces = [f for fl in face_lists for f in fl.faces]
    for i, f in enumerate(all_faces):
        f.idx = i
        f.count = 0
        f.material.wavelengths = np.asarray(wavelengths)
        f.max_length = max_length
    return all_faces


def singlet_c1(n_rays=10000, depth=4):
    wl = 0.55
    air = NondispersiveCurve(1.0)
    bk7 = glass(N_BK7)
    front = FullDielectricDispersiveMaterial(dispersion_inside=air, dispersion_outside=bk7,
                                             reflection_threshold=0.0, transmission_threshold=0.0)
    back = FullDielectricDispersiveMaterial(dispersion_inside=bk7, dispersion_outside=air,
                                            reflection_threshold=0.0, transmission_threshold=0.0)
    lens = FaceList()
    lens.faces = [SphericalFace(diameter=25.0, z_height=-2.5, curvature=-50.0, material=front),
                  SphericalFace(diameter=25.0, z_height=2.5, curvature=50.0, material=back)]
    n = sellmeier_n(N_BK7, wl)
    # thick-lens paraxial back focal length
    R1, R2, t = 50.0, -50.0, 5.0
    P = (n - 1) * (1 / R1 - 1 / R2 + (n - 1) * t / (n * R1 * R2))
    efl = 1.0 / P
    bfl = efl * (1 - (n - 1) * t / (n * R1))
    det = FaceList()
    det.faces = [CircularFace(diameter=60.0, z_plane=2.5 + bfl, material=OpaqueMaterial())]
    rays = hex_field(n_rays, 10.0, -10.0)
    info = {"n_glass": n, "efl_mm": efl, "bfl_mm": bfl, "detector_z_mm": 2.5 + bfl}
    return Workload("singlet_c1", rays, [wl], [lens, det], depth, 100.0, info)


def asphere_doublet_c2(n_rays=4194304, seed=12345, max_angle_deg=2.0, wl_range=(0.45, 0.65)):
    """The aspheric + achromat system: 4,194,304 rays on a 6 mm disk with
    angles U(+/-2 deg) and per-ray wavelengths U(0.45, 0.65) um by default;
    max_angle_deg=0 and wl_range=(0.55, 0.55) give its on-axis
    monochromatic bundle."""
    air = NondispersiveCurve(1.0)
    bk7, sf5 = glass(N_BK7), glass(N_SF5)
    mgf2 = NondispersiveCurve(1.38)
    t_coat = 0.55 / (4 * 1.38)
    shape = CircleShape(radius=10.0)
    # asphere: optical R = +30 mm (convex towards the source), k=-0.6, A4..A10
    # of the prescription in the usual sag convention z = c r^2/(1+sqrt(1-(1+k)
    # c^2 r^2)) + A4 r^4 + ...; built as raypier's PlanoAsphericLens builds its
    # A surface (lenses.py:466-477): AsphericFace curvature = -R, inverted
    # normals; planar back face at CT = 6 mm
    asph_mat = CoatedDispersiveMaterial(dispersion_inside=bk7, dispersion_outside=air, dispersion_coating=mgf2,
                                        coating_thickness=t_coat)
    A = {"A4": -2e-5, "A6": 1e-7, "A8": -3e-10, "A10": 5e-13}
    asphere = FaceList()
    asphere.faces = [AsphericFace(shape=shape, material=asph_mat, conic_const=-0.6, z_height=0.0, curvature=-30.0,
                                  A4=A["A4"], A6=A["A6"], A8=A["A8"], A10=A["A10"], invert_normals=True,
                                  atol=1e-16),
                     ShapedPlanarFace(shape=shape, material=asph_mat, z_height=6.0)]
    zA, ct1, ct2 = 16.0, 6.0, 3.0
    m1 = CoatedDispersiveMaterial(dispersion_inside=air, dispersion_outside=bk7, dispersion_coating=mgf2,
                                  coating_thickness=t_coat)
    m2 = CoatedDispersiveMaterial(dispersion_inside=bk7, dispersion_outside=sf5, dispersion_coating=mgf2,
                                  coating_thickness=0.0)
    m3 = CoatedDispersiveMaterial(dispersion_inside=sf5, dispersion_outside=air, dispersion_coating=mgf2,
                                  coating_thickness=t_coat)
    doublet = FaceList()
    doublet.faces = [SphericalFace(diameter=25.0, z_height=zA, curvature=-61.5, material=m1),
                     SphericalFace(diameter=25.0, z_height=zA + ct1, curvature=44.2, material=m2),
                     SphericalFace(diameter=25.0, z_height=zA + ct1 + ct2, curvature=128.2, material=m3)]
    # paraxial focus at 0.55 um by a y-nu trace of the five refracting surfaces
    wl0 = 0.55
    nb, nf = sellmeier_n(N_BK7, wl0), sellmeier_n(N_SF5, wl0)
    surfs = [(0.0, 30.0, 1.0, nb), (6.0, 1e300, nb, 1.0), (zA, 61.5, 1.0, nb), (zA + ct1, -44.2, nb, nf),
             (zA + ct1 + ct2, -128.2, nf, 1.0)]
    y, u, z_prev = 1.0, 0.0, 0.0
    for z, R, n1, n2 in surfs:
        y = y + u * (z - z_prev)
        u = (n1 * u - y * (n2 - n1) / R) / n2
        z_prev = z
    focus = z_prev - y / u
    det = FaceList()
    det.faces = [CircularFace(diameter=80.0, z_plane=focus, material=OpaqueMaterial())]
    rays, wl = disk_field(n_rays, 6.0, -5.0, max_angle_deg, wl_range, seed)
    info = {"asphere_R_mm": 30.0, "asphere_CT_mm": 6.0, "doublet_z_mm": zA, "CT1_mm": ct1, "CT2_mm": ct2,
            "coating_um": t_coat, "focus_z_mm": focus, "seed": seed}
    return Workload("asphere_doublet_c2", rays, wl, [asphere, doublet, det], 10, 200.0, info)


def _sag(R, h):
    """Sag of a sphere of optical radius R at height h (sign of R)."""
    if abs(R) > 1e8:
        return 0.0
    return R - math.copysign(math.sqrt(R * R - h * h), R)


def device_disk_rays(n, radius, z0, n_wavelengths, seed):
    """Collimated rays (+z, E along x, E1 = 1) uniform on a disk, built in HBM
    with a seeded torch generator: ray_ident = index, wavelength index
    uniform over the table."""
    import torch
    from ._device import DeviceRays
    from ._abi import F64_FIELDS, U32_FIELDS
    g = torch.Generator(device="cuda")
    g.manual_seed(int(seed))
    dev = DeviceRays(max(n, 1))
    f, u = dev.f64(), dev.u32()
    f.zero_()
    rr = radius * torch.sqrt(torch.rand(n, generator=g, device="cuda", dtype=torch.float64))
    th = 2 * math.pi * torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
    k = F64_FIELDS.index
    f[k("ox"), :n] = rr * torch.cos(th)
    f[k("oy"), :n] = rr * torch.sin(th)
    f[k("oz"), :n] = z0
    f[k("dz"), :n] = 1.0
    f[k("ex"), :n] = 1.0
    f[k("n_re"), :n] = 1.0
    f[k("e1_re"), :n] = 1.0
    f[k("length"), :n] = float("inf")
    q = U32_FIELDS.index
    u[q("wavelength_idx"), :n] = torch.randint(0, n_wavelengths, (n,), generator=g, device="cuda",
                                               dtype=torch.int32)
    u[q("parent_idx"), :n] = -1
    u[q("end_face_idx"), :n] = -1
    u[q("ray_ident"), :n] = torch.arange(n, device="cuda", dtype=torch.int32)
    u[q("ray_type_id"), :n] = 0
    return RayCollection.from_device(dev, n)


def lens_job_c5(seed, n_rays=None, device_rays=False):
    """One serving job: a random prescription of 2-8 refracting surfaces
    (1-4 singlets) drawn from sphere, conic and even-asphere (A4..A10)
    types, curvatures U(+/-1/20) mm^-1, thicknesses and air gaps U(2, 10) mm
    (raised where an edge would be thinner than 0.5 mm), each singlet in
    N-BK7 or N-SF11, uncoated Fresnel interfaces with the default 0.1 power
    thresholds; collimated rays on a 4 mm-radius disk at three wavelengths
    (F, d, C lines); opaque detector at the paraxial focus (or 50 mm behind
    the last surface for a diverging system).  Ray count log-uniform in
    [1e3, 2e6] unless given.  Conic surfaces are AsphericFaces with zero
    polynomial terms."""
    rng = np.random.default_rng(seed)
    if n_rays is None:
        n_rays = int(round(math.exp(rng.uniform(math.log(1e3), math.log(2e6)))))
    n_elem = int(rng.integers(1, 5))
    air = NondispersiveCurve(1.0)
    h = 6.0                                   # clear aperture radius
    shape = CircleShape(radius=h)
    wl = np.array([0.4861327, 0.5875618, 0.6562725])
    faces, surfs, presc = [], [], []
    media = [air]
    z = 0.0
    prev_back_R, gap = None, 0.0
    for e in range(n_elem):
        coefs = N_BK7 if rng.random() < 0.5 else N_SF11
        g = glass(coefs)
        media.append(g)
        gi = len(media) - 1
        n_d = sellmeier_n(coefs, 0.5875618)
        Rs = []
        for s in range(2):
            c = rng.uniform(-1.0 / 20, 1.0 / 20)
            Rs.append(1.0 / c if abs(c) > 1e-9 else 1e9)
        if prev_back_R is not None:           # air gap: keep 0.5 mm at the edge
            z += max(0.0, 0.5 - gap + _sag(prev_back_R, h) - _sag(Rs[0], h))
        t = rng.uniform(2.0, 10.0)
        t = max(t, 0.5 + _sag(Rs[0], h) - _sag(Rs[1], h))
        for s, R in enumerate(Rs):
            kind = rng.choice(["sphere", "conic", "asphere"])
            zs = z if s == 0 else z + t
            front = s == 0
            # sphere: outside = +z side; asphere (inverted normals): outside = -z side
            if kind == "sphere":
                mat = FullDielectricDispersiveMaterial(dispersion_inside=air if front else g,
                                                       dispersion_outside=g if front else air)
                f = SphericalFace(diameter=2 * h, z_height=zs, curvature=-R, material=mat)
            else:
                k = float(rng.uniform(-1.0, 0.5))
                A = [0.0] * 4
                if kind == "asphere":
                    A = [rng.normal(0, 1e-5), rng.normal(0, 1e-7), rng.normal(0, 1e-10), rng.normal(0, 1e-13)]
                mat = FullDielectricDispersiveMaterial(dispersion_inside=g if front else air,
                                                       dispersion_outside=air if front else g)
                f = AsphericFace(shape=shape, material=mat, conic_const=k, z_height=zs, curvature=-R,
                                 A4=A[0], A6=A[1], A8=A[2], A10=A[3], invert_normals=True, atol=1e-16)
            faces.append(f)
            surfs.append((zs, R, 1.0 if front else n_d, n_d if front else 1.0))
            presc.append({"kind": str(kind), "z": zs, "c": 1.0 / R, "k": 0.0 if kind == "sphere" else k,
                          "A": [0.0] * 4 if kind == "sphere" else [float(a) for a in A], "aperture": h,
                          "before": 0 if front else gi, "after": gi if front else 0})
        prev_back_R = Rs[1]
        gap = rng.uniform(2.0, 10.0)          # air gap to the next element
        z += t + gap
    last_z = surfs[-1][0]
    y, u, z_prev = 1.0, 0.0, surfs[0][0]
    for zs, R, n1, n2 in surfs:
        y = y + u * (zs - z_prev)
        u = (n1 * u - y * (n2 - n1) / R) / n2
        z_prev = zs
    bfl = -y / u if u < 0 else float("inf")
    det_z = last_z + (bfl if 1.0 < bfl < 500.0 else 50.0)
    lens = FaceList()
    lens.faces = faces
    det = FaceList()
    det.faces = [CircularFace(diameter=1000.0, z_plane=det_z, material=OpaqueMaterial())]
    z0 = surfs[0][0] - 5.0
    info = {"surfaces": len(faces), "detector_z_mm": det_z, "detector_face": len(faces), "seed": seed,
            "n_rays": n_rays, "prescription": presc, "media": media}
    if device_rays:
        return Workload("lens_job_c5", None, wl, [lens, det], len(faces) + 3, 1000.0, info,
                        source=lambda: device_disk_rays(n_rays, 4.0, z0, 3, seed))
    r = _disk_collimated(n_rays, 4.0, z0, rng)
    r["wavelength_idx"] = rng.integers(0, 3, n_rays).astype(np.uint32)
    return Workload("lens_job_c5", r, wl, [lens, det], len(faces) + 3, 1000.0, info)


def _disk_collimated(n, radius, z0, rng):
    rr = radius * np.sqrt(rng.random(n))
    th = 2 * np.pi * rng.random(n)
    r = base_rays(n)
    r["origin"][:, 0] = rr * np.cos(th)
    r["origin"][:, 1] = rr * np.sin(th)
    r["origin"][:, 2] = z0
    return r


def collimated_gausslets(n_gausslets, radius, beam_waist, wavelength, z0, blending=1.5, beam_power=1.0,
                         E_vector=(1.0, 0.0, 0.0), E1_amp=1.0 + 0j, E2_amp=0j):
    """Collimated Gaussian beam along +z as a grid of gausslets, built the
    way raypier's CollimatedGaussletSource builds it
    (gausslet_sources.py:386-457): a lattice with rows skewed by 30 degrees,
    points selected by the same radial rule, Gaussian amplitude weights,
    power scaling, and parabasal rays of radius spacing / blending with the
    waist at the base origin.  The lattice spacing is chosen so that the
    selection holds at least n_gausslets points; the n_gausslets nearest the
    axis by that rule (index order on ties) are kept."""
    from .ctracer import GaussletCollection
    cosV, sinV = math.cos(math.pi * 30 / 180.), math.sin(math.pi * 30 / 180.)

    def lattice(resolution):
        spacing = radius / resolution
        nsteps = int(radius / (spacing * cosV))
        i = np.arange(-nsteps - 1, nsteps + 2)
        vi, vj = np.meshgrid(i, i)
        vi, vj = vi.ravel(), vj.ravel()
        ri = ((vi + sinV * vj) ** 2 + (cosV * vj ** 2))
        sel = ri < (radius / spacing) ** 2
        return spacing, vi[sel], vj[sel], ri[sel]

    res = math.sqrt(n_gausslets * cosV / math.pi)
    while True:
        spacing, xi, yj, ri = lattice(res)
        if len(ri) >= n_gausslets:
            break
        res *= 1.01
    keep = np.sort(np.argsort(ri, kind="stable")[:n_gausslets])
    xi, yj, ri = xi[keep], yj[keep], ri[keep]
    direction = np.array([0.0, 0.0, 1.0])
    v = np.array([1.0, 0.0, 0.0])
    d1 = np.cross(direction, v)
    d1 /= np.linalg.norm(d1)
    d2 = np.cross(direction, d1)
    d2 /= np.linalg.norm(d2)
    d2 = cosV * d2 + sinV * d1
    E = np.cross(np.cross(np.asarray(E_vector, float), direction), direction)
    E /= np.linalg.norm(E)
    offsets = (xi[:, None] * d1 + yj[:, None] * d2) * spacing
    gauss = np.exp(-(ri * (spacing ** 2) / (beam_waist ** 2)))
    P1 = (E1_amp * np.conj(E1_amp)).real
    P2 = (E2_amp * np.conj(E1_amp)).real
    scaling = math.sqrt(beam_power / ((P1 + P2) * (gauss ** 2)).sum())
    r = base_rays(len(xi))
    r["origin"] = offsets + np.array([0.0, 0.0, z0])
    r["direction"] = direction
    r["E_vector"] = E
    r["E1_amp"] = scaling * E1_amp * gauss
    r["E2_amp"] = scaling * E2_amp * gauss
    r["normal"] = (0.0, 1.0, 0.0)
    r["ray_type_id"] = 1 << 1
    r["length"] = np.inf
    g = GaussletCollection.from_rays(r)
    wl = np.array([wavelength])
    g.wavelengths = wl
    g.config_parabasal_rays(wl, spacing / blending, 0.0)
    return g, spacing


def gausslet_c4(n_gausslets=262144, grid=512, span_mm=0.040):
    """Gausslet field workload: a collimated Gaussian beam (lambda = 1.064 um,
    1/e^2 field radius 2 mm, sampled to 4 mm) from z = -80 mm through an
    N-BK7 plano-convex asphere (optical R = f (n - 1) for f = 20 mm, i.e.
    NA 0.1 at the beam radius; k = -0.5, A4 = 1e-5; CT 4 mm), uncoated;
    the gausslets leaving the back face are summed on a grid x grid plane
    spanning span_mm around the least-squares focus of their base rays.
    Optical path to the focus ~1e5 wavelengths."""
    wl = 1.064
    n = sellmeier_n(N_BK7, wl)
    R = 20.0 * (n - 1.0)
    g_bk7 = glass(N_BK7)
    air = NondispersiveCurve(1.0)
    shape = CircleShape(radius=6.0)
    mat = FullDielectricDispersiveMaterial(dispersion_inside=g_bk7, dispersion_outside=air)
    lens = FaceList()
    lens.faces = [AsphericFace(shape=shape, material=mat, conic_const=-0.5, z_height=0.0, curvature=-R, A4=1e-5,
                               invert_normals=True, atol=1e-16),
                  ShapedPlanarFace(shape=shape, material=mat, z_height=4.0)]
    gc, spacing = collimated_gausslets(n_gausslets, 4.0, 2.0, wl, -80.0)
    info = {"n_gausslets": n_gausslets, "spacing_mm": spacing, "R_mm": R, "grid": grid, "span_mm": span_mm}
    return gc, [lens], info


def field_grid(focus, grid, span_mm):
    """grid x grid points in the z = focus[2] plane, centred on focus."""
    g = np.linspace(-span_mm / 2, span_mm / 2, grid)
    gx, gy = np.meshgrid(g + focus[0], g + focus[1])
    return np.column_stack([gx.ravel(), gy.ravel(), np.full(gx.size, focus[2])])


def noll_to_nm(j):
    """(n, m) of Noll index j >= 1 (cosine terms at even j, sine at odd)."""
    n = int((-1 + math.sqrt(8 * (j - 1) + 1)) / 2)
    p = j - n * (n + 1) // 2 - 1
    m = (n % 2) + 2 * ((p + (n + 1) % 2) // 2)
    return n, (m if j % 2 == 0 else -m)


def noll_to_ansi(j):
    n, m = noll_to_nm(j)
    return (n * (n + 2) + m) // 2


def _placed(faces, columns, translation):
    """FaceList whose local axes map to the given global columns (x', y',
    z') and whose local origin sits at `translation`."""
    from .ctracer import Transform
    R = np.array(columns, dtype=np.float64).T          # columns -> matrix
    fl = FaceList()
    fl.faces = faces
    t = np.asarray(translation, dtype=np.float64)
    fl.transform = Transform(rotation=R.tolist(), translation=t.tolist())
    fl.inverse_transform = Transform(rotation=R.T.tolist(), translation=(-(R.T @ t)).tolist())
    return fl


def michelson_c3(n_rays=1048576, order=1, seed=2025, depth=12, threshold=1e-4, wavelength=0.6328, orders=None):
    """Michelson interferometer in the x-y plane, light along +x:
      * 25 mm N-BK7 beamsplitter cube at the origin -- four extruded side
        faces and two polygon end faces with FullDielectricMaterial (relative
        reflection / transmission thresholds `threshold`), the diagonal
        (-a,-a)-(a,a) a PartiallyReflectiveMaterial(0.5);
      * arm +x: flat PEC mirror (radius 15 mm) at x = 50;
      * arm +y: PEC mirror at y = 50 with a Zernike distortion, Noll terms
        4-15 (ANSI-indexed), coefficients N(0, 0.05 um), unit radius 15 mm;
      * output port -y: 600 lines/mm reflective grating at y = -50 tilted by
        10 degrees about z, one diffraction order `order` per trace as the
        reference's grating material has, or with orders=(-1, 0, 1) this
        build's multi-order grating: up to three children per hit, each
        order's child as a one-order grating emits it;
      * an opaque 200 mm box around everything.
    Rays: n_rays collimated along +x from x = -40 on a 4 mm disk, random
    elliptical polarisation (E1 = cos(chi) e^{i phi1}, E2 = sin(chi) e^{i phi2}).
    The trace runs to `depth` generations."""
    from .cfaces import ExtrudedPlanarFace, PolygonFace, RectangularFace, DistortionFace
    from .cmaterials import FullDielectricMaterial, PartiallyReflectiveMaterial, PECMaterial, \
        DiffractionGratingMaterial
    from .cdistortions import ZernikeDistortion
    rng = np.random.default_rng(seed)
    a, n_glass = 12.5, sellmeier_n(N_BK7, wavelength)

    def glass_mat(n_in, n_out):
        return FullDielectricMaterial(n_inside=n_in, n_outside=n_out, reflection_threshold=threshold,
                                      transmission_threshold=threshold)
    sq = [(-a, a), (a, a), (a, -a), (-a, -a)]                       # clockwise: outward normals
    cube_faces = []
    for (x1, y1), (x2, y2) in zip(sq, sq[1:] + sq[:1]):
        cube_faces.append(ExtrudedPlanarFace(x1=x1, y1=y1, x2=x2, y2=y2, z1=-a, z2=a,
                                             material=glass_mat(n_glass, 1.0)))
    cube_faces.append(PolygonFace(z_plane=a, xy_points=sq, material=glass_mat(1.0, n_glass)))
    cube_faces.append(PolygonFace(z_plane=-a, xy_points=sq, material=glass_mat(n_glass, 1.0)))
    cube_faces.append(ExtrudedPlanarFace(x1=-a, y1=-a, x2=a, y2=a, z1=-a, z2=a,
                                         material=PartiallyReflectiveMaterial(reflectivity=0.5)))
    cube = FaceList()
    cube.faces = cube_faces
    mirror1 = _placed([ShapedPlanarFace(shape=CircleShape(radius=15.0), z_height=0.0, material=PECMaterial())],
                      [(0, 1, 0), (0, 0, 1), (1, 0, 0)], (50.0, 0.0, 0.0))
    noll = list(range(4, 16))
    coefs = rng.normal(0.0, 0.05e-3, len(noll))                    # 0.05 um in mm
    zd = ZernikeDistortion([(noll_to_ansi(j), c) for j, c in zip(noll, coefs)], unit_radius=15.0)
    base = ShapedPlanarFace(shape=CircleShape(radius=15.0), z_height=0.0)
    mirror2 = _placed([DistortionFace(base_face=base, distortion=zd, material=PECMaterial())],
                      [(0, 0, 1), (1, 0, 0), (0, 1, 0)], (0.0, 50.0, 0.0))
    th = math.radians(10.0)
    ex = (math.cos(th), math.sin(th), 0.0)
    grating = _placed([ShapedPlanarFace(shape=CircleShape(radius=20.0), z_height=0.0,
                                        material=DiffractionGratingMaterial(lines_per_mm=600, order=order,
                                                                           efficiency=0.33, orders=orders,
                                                                           origin=(0.0, -50.0, 0.0)))],
                      [ex, (0, 0, 1), (math.sin(th), -math.cos(th), 0.0)], (0.0, -50.0, 0.0))
    b = 100.0
    box_sq = [(-b, b), (b, b), (b, -b), (-b, -b)]
    box_faces = [ExtrudedPlanarFace(x1=x1, y1=y1, x2=x2, y2=y2, z1=-b, z2=b, material=OpaqueMaterial())
                 for (x1, y1), (x2, y2) in zip(box_sq, box_sq[1:] + box_sq[:1])]
    box_faces += [RectangularFace(z_plane=b, length=2 * b, width=2 * b, material=OpaqueMaterial()),
                  RectangularFace(z_plane=-b, length=2 * b, width=2 * b, material=OpaqueMaterial())]
    box = FaceList()
    box.faces = box_faces
    # rays
    rr = 4.0 * np.sqrt(rng.random(n_rays))
    ph = 2 * np.pi * rng.random(n_rays)
    r = base_rays(n_rays)
    r["origin"][:, 0] = -40.0
    r["origin"][:, 1] = rr * np.cos(ph)
    r["origin"][:, 2] = rr * np.sin(ph)
    r["direction"] = (1.0, 0.0, 0.0)
    psi = 2 * np.pi * rng.random(n_rays)
    r["E_vector"] = np.stack([np.zeros(n_rays), np.cos(psi), np.sin(psi)], axis=1)
    chi = 0.5 * np.pi * rng.random(n_rays)
    phi1, phi2 = 2 * np.pi * rng.random(n_rays), 2 * np.pi * rng.random(n_rays)
    r["E1_amp"] = np.cos(chi) * np.exp(1j * phi1)
    r["E2_amp"] = np.sin(chi) * np.exp(1j * phi2)
    info = {"order": order, "orders": orders, "zernike_noll_4_15_mm": [float(c) for c in coefs], "n_glass": n_glass,
            "threshold": threshold, "seed": seed}
    return Workload("michelson_c3", r, [wavelength], [cube, mirror1, mirror2, grating, box], depth, 500.0, info)
