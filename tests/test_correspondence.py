"""Correspondence search and its filters, each provoked on purpose, in the
manner of the reference's tests/test_correspondence.py."""

from __future__ import annotations

import numpy as np

ACCEPTED, DISTANCE, NORMAL, SLOPE, MASKED = 0, 1, 2, 3, 4


def square():
    """The unit square at z = 0 as two triangles, normal +z."""
    pts = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.0, 1.0, 0], [0, 1.0, 0]])
    return pts, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int64)


UP = np.array([0.0, 0.0, 1.0])


def test_every_filter_rejects_for_its_own_reason(impl):
    pts, faces = square()
    q = np.array([[0.45, 0.55, 0.08],    # hovering above: accepted
                  [0.5, 0.3, 6.0],       # beyond the distance threshold
                  [0.7, 0.4, 0.02],      # normal points the other way
                  [1.4, 0.5, 0.0],       # off the edge, sliding
                  [0.2, 0.8, 0.003]])    # masked
    n = np.array([UP, UP, -UP, UP, UP])
    mask = np.array([1, 1, 1, 1, 0])
    r = impl.find_correspondences(pts, faces, q, n, distance_threshold=1.0, min_normal_cosine=0.5, min_slope=0.5,
                                  mask=mask)
    assert list(np.asarray(r["rejection"])) == [ACCEPTED, DISTANCE, NORMAL, SLOPE, MASKED]
    c = r["counts"]
    assert (c["total"], c["accepted"], c["rejected_distance"], c["rejected_normal"], c["rejected_slope"],
            c["rejected_masked"]) == (5, 1, 1, 1, 1, 1)


def test_the_slope_threshold_is_a_minimum(impl):
    pts, faces = square()
    assert impl.find_correspondences(pts, faces, np.array([[0.4, 0.4, 0.3]]), UP[None],
                                     min_slope=0.5)["counts"]["accepted"] == 1
    sliding = np.array([[-0.25, 0.4, 0.0]])
    assert impl.find_correspondences(pts, faces, sliding, UP[None], min_slope=0.5)["counts"]["rejected_slope"] == 1
    assert impl.find_correspondences(pts, faces, sliding, UP[None], min_slope=0.0)["counts"]["accepted"] == 1


def test_statistics_are_taken_over_accepted_rows_only(impl):
    pts, faces = square()
    h = np.array([0.015, 0.025, 0.045, 3.0])
    q = np.stack([np.full(4, 0.4), np.full(4, 0.6), h], 1)
    r = impl.find_correspondences(pts, faces, q, np.repeat(UP[None], 4, 0), distance_threshold=1.0)
    c = r["counts"]
    assert c["accepted"] == 3 and c["rejected_distance"] == 1
    assert c["distance_mean"] == np.mean(h[:3])
    assert c["distance_max"] == h[2]
    # p95 of three accepted distances: sorted[floor(0.95 * 2)] = the second.
    assert c["distance_p95"] == h[1]


def test_a_query_on_the_surface_passes_a_tight_slope(impl):
    pts, faces = square()
    r = impl.find_correspondences(pts, faces, np.array([[0.5, 0.25, 0.0]]), UP[None], min_slope=1e-6)
    assert r["counts"]["accepted"] == 1 and np.asarray(r["distances"])[0] == 0.0


def test_the_correspondence_lies_on_the_surface(impl):
    pts, faces = square()
    r = impl.find_correspondences(pts, faces, np.array([[0.35, 0.65, 0.07]]), UP[None])
    np.testing.assert_allclose(np.asarray(r["points"])[0], [0.35, 0.65, 0.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.asarray(r["normals"])[0], UP, rtol=0, atol=1e-12)
