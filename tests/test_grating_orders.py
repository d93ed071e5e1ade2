"""The multi-order grating (DiffractionGratingMaterial(orders=...), not in
the reference, whose material has one order): the oracle's children of a
multi-order hit are exactly the children one-order gratings of the listed
orders emit, in list order, with evanescent orders dropped; and a C3 trace
with orders (-1, 0, +1) has up to three children per grating hit."""
import numpy as np

import oracle
from raypier_optics_amd import workloads, _abi
from raypier_optics_amd.cmaterials import DiffractionGratingMaterial
from raypier_optics_amd.cfaces import ShapedPlanarFace
from raypier_optics_amd.cshapes import CircleShape
from raypier_optics_amd.ctracer import FaceList
from raypier_optics_amd._scene import compile_scene


def _scene(mats):
    fl = FaceList()
    fl.faces = [ShapedPlanarFace(shape=CircleShape(radius=20.0), z_height=float(k), material=m)
                for k, m in enumerate(mats)]
    workloads.prepare_faces([fl], [0.6328, 1.55], 100.0)
    return compile_scene([fl])


def _ray(direction, wl_idx=0):
    r = workloads.base_rays(1)
    r["origin"] = (0.3, -0.2, -5.0)
    r["direction"] = direction / np.linalg.norm(direction)
    r["E_vector"] = (1.0, 0.0, 0.0)
    r["E1_amp"], r["E2_amp"] = 0.6 + 0.2j, -0.3 + 0.7j
    r["wavelength_idx"] = wl_idx
    r["refractive_index"] = 1.0
    return r


def test_record_packing():
    m = DiffractionGratingMaterial(lines_per_mm=600, orders=(-1, 0, 2), efficiency=0.3)
    t = _scene([m])
    rec = t.materials[0]
    assert rec["order"] == -1 and rec["n_orders"] == 3
    assert np.int8(rec["more_orders"] & 0xFF) == 0 and np.int8((rec["more_orders"] >> 8) & 0xFF) == 2
    assert t.max_children == 3
    one = _scene([DiffractionGratingMaterial(lines_per_mm=600, order=-1)])
    assert one.materials[0]["n_orders"] == 0 and one.max_children == 1


def test_children_are_the_one_order_children():
    rng = np.random.default_rng(4)
    orders = (-1, 0, 1)
    for lpm in (300.0, 600.0, 1800.0):
        multi = DiffractionGratingMaterial(lines_per_mm=lpm, orders=orders, efficiency=0.33, origin=(1.0, 2.0, 0.0))
        singles = [DiffractionGratingMaterial(lines_per_mm=lpm, order=o, efficiency=0.33, origin=(1.0, 2.0, 0.0))
                   for o in orders]
        t = _scene([multi] + singles)
        for trial in range(40):
            d = np.array([rng.normal(0, 0.5), rng.normal(0, 0.5), 1.0])
            wl_idx = trial % 2
            ray = _ray(d, wl_idx)
            point = np.array([0.3, -0.2, 0.0]) + rng.normal(0, 0.1, 3) * [1, 1, 0]
            normal, tangent = np.array([0.0, 0.0, -1.0]), np.array([1.0, 0.0, 0.0])
            got = oracle.eval_child_ray(t, 0, ray, 7, point, normal, tangent, [0.6328, 1.55])
            want = [oracle.eval_child_ray(t, 1 + k, ray, 7, point, normal, tangent, [0.6328, 1.55])
                    for k in range(3)]
            want = np.concatenate(want)
            assert len(got) == len(want) <= 3
            assert got.tobytes() == want.tobytes()
            if lpm == 1800.0 and wl_idx == 1:
                assert len(got) < 3        # 1.55 um at 1800 l/mm: the first orders are evanescent


def test_c3_multi_order_trace():
    w = workloads.michelson_c3(n_rays=600, orders=(-1, 0, 1))
    workloads.prepare_faces(w.face_lists, w.wavelengths, w.max_length)
    t = compile_scene(w.face_lists)
    assert t.max_children == 3
    traced, hits = oracle.trace_rays(t, w.rays, w.max_length, w.wavelengths, w.recursion_limit)
    grating = [k for k, f in enumerate(t.faces) if f["material"] >= 0 and
               t.materials[f["material"]]["type"] == _abi.MAT_GRATING][0]
    gidx = int(t.faces[grating]["global_idx"])
    triples = 0
    for parent, child in zip(traced[:-1], traced[1:]):
        on_grating = np.nonzero(parent["end_face_idx"] == gidx)[0]
        if len(on_grating) == 0 or len(child) == 0:
            continue
        counts = np.bincount(child["parent_idx"], minlength=len(parent))[on_grating]
        assert counts.max() <= 3
        triples += int((counts == 3).sum())
    assert triples > 0
