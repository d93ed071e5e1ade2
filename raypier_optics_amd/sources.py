"""Ray-source helpers that work on traced generations.

`set_face_sequence` is BaseRaySource.set_face_sequence (sources.py:138-155):
the face sequence for sequential tracing (tracer.trace_ray_sequence), taken
as the most common end face of each traced generation.  The per-generation
end-face histogram is one GPU launch over all generations
(rpx_face_histogram); the vote itself is read off the histogram.
"""
import numpy as np

from . import _abi
from ._lib import load, check, require_gpu, stream_handle
from .results import generation_table

NO_FACE = 0xFFFFFFFF


def end_face_histogram(traced_rays, n_bins):
    """uint64 (n_generations, n_bins + 2): counts of end_face_idx == b for
    b < n_bins, then of 0xFFFFFFFF, then of every other value."""
    torch = require_gpu()
    lib = load()
    gens, devs = generation_table(traced_rays)
    ws = torch.empty(int(lib.rpx_lineage_workspace_bytes(len(traced_rays), 0)), dtype=torch.uint8, device="cuda")
    counts = torch.empty((len(traced_rays), n_bins + 2), dtype=torch.int64, device="cuda")
    check(lib.rpx_face_histogram(gens, len(traced_rays), n_bins, counts.data_ptr(), ws.data_ptr(), ws.numel(),
                                 stream_handle()))
    return counts.cpu().numpy().astype(np.uint64)


def _most_common(counts, n_bins, rays):
    """The value numpy.unique(end_face_idx, return_counts=True) puts first
    among those with the largest count (the smallest such value)."""
    if counts.sum() == 0:
        raise ValueError("attempt to get argmax of an empty sequence")
    if counts[n_bins + 1]:
        # values other than 0..n_bins-1 and 0xFFFFFFFF: count them exactly
        ids, c = np.unique(rays.base_rays.end_face_idx, return_counts=True)
        return int(ids[c.argmax()])
    k = int(np.argmax(counts[: n_bins + 1]))
    return k if k < n_bins else NO_FACE


def set_face_sequence(traced_rays, face_lists, all_faces):
    """[(FaceList, face index in it), ...] for the generations of
    traced_rays, stopping at the first generation whose most common end face
    is not a face of face_lists."""
    global_map = {face: i for i, face in enumerate(all_faces)}
    face_map = {global_map[face]: (fl, face_idx) for fl in face_lists for face_idx, face in enumerate(fl.faces)}
    if not traced_rays:
        return []
    n_bins = len(all_faces)
    if n_bins > 4094:
        raise ValueError("more than 4094 faces")
    counts = end_face_histogram(traced_rays, n_bins)
    face_sequence = []
    for g, rays in enumerate(traced_rays):
        most_common = _most_common(counts[g], n_bins, rays)
        try:
            face_tuple = face_map[most_common]
        except KeyError:
            break
        face_sequence.append(face_tuple)
    return face_sequence
