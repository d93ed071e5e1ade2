"""Generation drivers (drop-in for raypier.core.tracer)."""
from itertools import chain

import numpy

from .ctracer import RayCollection, GaussletCollection, _TraceContext, _generation


def _prepare(input_rays, face_lists, max_length):
    input_rays.reset_length(max_length)
    wavelengths = numpy.asarray(input_rays.wavelengths)
    all_faces = list(chain(*(fs.faces for fs in face_lists)))
    for i, f in enumerate(all_faces):
        f.idx = i
        f.count = 0
        f.update()
        f.material.wavelengths = wavelengths
        f.max_length = max_length
    return all_faces


def trace_rays(input_rays, face_lists, recursion_limit=100, max_length=100.0):
    """Non-sequential trace of a RayCollection or GaussletCollection
    (core/tracer.py:9-47): returns (traced_rays, all_faces).

    As in the reference, each collection is appended before it is traced, so
    with recursion_limit=L the L-th traced generation is not returned (its
    parents still receive their length and end_face_idx)."""
    all_faces = _prepare(input_rays, face_lists, max_length)
    face_sets = list(face_lists)
    for fs in face_sets:
        fs.sync_transforms()
    ctx = _TraceContext(face_sets, all_faces)
    decomp_faces = [f for f in all_faces if f.material.is_decomp_material()]
    traced_rays = []
    rays = input_rays
    count = 0
    while rays.n_rays > 0 and count < recursion_limit:
        traced_rays.append(rays)
        rays = _generation(ctx, rays, max_length, flush=False, decomp_faces=decomp_faces)
        count += 1
    ctx.flush_counts()
    return traced_rays, all_faces


def trace_ray_sequence(input_rays, face_sequence, recursion_limit=100, max_length=100.0):
    """Sequential trace through (FaceList, face_idx) pairs (core/tracer.py:50-99)."""
    input_rays.reset_length(max_length)
    traced_rays = [input_rays]
    face_lists = [fl for fl, fidx in face_sequence]
    face_idx_list = [fidx for fl, fidx in face_sequence]
    all_faces = _prepare(input_rays, face_lists, max_length)
    decomp_faces = [f for f in all_faces if f.material.is_decomp_material()]
    contexts = {}
    rays = input_rays
    count = 0
    for fl, fidx in zip(face_lists, face_idx_list):
        fl.sync_transforms()
        key = id(fl)
        if key not in contexts:
            contexts[key] = _TraceContext([fl], all_faces)
        rays = _generation(contexts[key], rays, max_length, 0, fidx, decomp_faces=decomp_faces)
        if (count > recursion_limit) or (rays.n_rays == 0):
            break
        traced_rays.append(rays)
        count += 1
    return traced_rays, all_faces
