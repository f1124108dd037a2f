"""Adaptive (AMR) meshes and per-step numerics against the reference.

The leaf-mesh builder is host code of the product (torch_amd.torchAdaptiveGeometry)
and is checked against the meshes the reference built (tests/golden/adaptive.npz,
made by make_adaptive_fixtures.py): keys, face lists, areas, offsets, volumes,
divergences and gradient matrices. The oracle's per-step numerics
(oracle/adaptive.py) are then checked on those meshes against the reference's
rhs (both orders), cfl and SSPRK2 advance.
"""

import json

import numpy as np
import pytest

from conftest import golden
from oracle import adaptive as oa

INTEGER = ("keys", "left", "right", "directions", "sides")
FLOAT = ("face_areas", "offset_left", "offset_right", "volumes", "divergence", "displacement", "weights", "inverse",
         "centers", "widths")


def amr_config(spec):
    from torch_amd import torchAdaptiveGeometry as tag
    from torch_amd.torchGeometry import session_faces

    dim = len(spec["counts"])
    c = dict(dimension=dim, counts_xyz=spec["counts"], lengths_xyz=spec["lengths"], boundaries_xyz=["Open"] * dim,
             origins_xyz=spec["origins"], geometry=spec["geometry"], t_end=0.02, cfl=0.4, gamma=1.4,
             limiter="VanLeer", reactive=spec["reactive"], kinetic_model="arrhenius",
             inlet_primitive=[1.0, 0.0, 100.0, 101325.0], max_steps=100000)
    c["face_boundaries"] = session_faces(c, None)
    c["amr"] = tag.settings(**spec["amr"])
    return c


@pytest.fixture(scope="module")
def d():
    return golden("adaptive")


def mesh_of(d, tag):
    from torch_amd.torchAdaptiveGeometry import LeafMesh

    c = amr_config(json.loads(str(d[tag + "_spec"])))
    return LeafMesh(c, d[tag + "_keys"]), c


def test_leaf_meshes_match_the_reference(d):
    for tag in d["mesh_cases"]:
        mesh, _ = mesh_of(d, tag)
        for f in INTEGER:
            np.testing.assert_array_equal(getattr(mesh, f), d[f"{tag}_{f}"], err_msg=f"{tag} {f}")
        for f in FLOAT:
            np.testing.assert_allclose(getattr(mesh, f), d[f"{tag}_{f}"], rtol=1e-14, atol=1e-15,
                                       err_msg=f"{tag} {f}")


def close(a, b, rel=1e-12, what=""):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    scale = np.max(np.abs(b), axis=tuple(range(1, b.ndim)), keepdims=True) if b.ndim > 1 else np.max(np.abs(b))
    err = float(np.max(np.abs(a - b) / np.maximum(scale, 1e-300))) if b.size else 0.0
    assert err <= rel, f"{what}: {err:.3e}"


def test_rhs_cfl_and_advance_match_the_reference(d):
    for tag in d["mesh_cases"]:
        mesh, c = mesh_of(d, tag)
        state = d[tag + "_state"]
        fr = [d[tag + "_lam"]] if tag + "_lam" in d else []
        for order, first in (("2", False), ("1", True)):
            change, species, boundary, source = oa.rhs(mesh, state, fr, c, first)
            close(change, d[f"{tag}_change{order}"], what=f"{tag} change{order}")
            assert np.max(np.abs(boundary - d[f"{tag}_boundary{order}"])) <= 1e-12 * max(
                np.max(np.abs(d[f"{tag}_boundary{order}"])), 1.0), tag
            close(source, d[f"{tag}_source{order}"], what=f"{tag} source{order}")
            if fr:
                close(species[0], d[f"{tag}_species{order}"], what=f"{tag} species{order}")
        dt = oa.cfl(mesh, state, fr, c)
        assert abs(dt - float(d[tag + "_dt"])) <= 1e-12 * dt
        final, qs, b, src, _ = oa.advance(mesh, state, fr, 0.5 * dt, c)
        close(final, d[tag + "_adv_state"], what=f"{tag} advance")
        if fr:
            close(qs[0], d[tag + "_adv_lam"], what=f"{tag} advance lam")


def test_settings_and_tree_validation():
    from torch_amd.torchAdaptiveGeometry import LeafMesh, children, settings

    for options in ({"max_level": 9}, {"max_cells": True}, {"refine_threshold": 0.01, "coarsen_threshold": 0.02},
                    {"bogus": 1}):
        with pytest.raises(ValueError):
            settings(**options)
    c = amr_config(dict(counts=[8], lengths=[1.0], origins=[0.0], geometry="cartesian", reactive=False,
                        amr=dict(max_level=2)))
    root = LeafMesh(c).keys
    for keys in (root[:-1], np.vstack((root, root[0])), np.vstack((root, children(tuple(root[0])))),
                 root.astype(float)):
        with pytest.raises(ValueError):
            LeafMesh(c, keys)
    keys = np.vstack((root[1:], children(tuple(root[0]))))
    fine = keys[(keys[:, 0] == 1) & (keys[:, 1] == 1)][0]
    keys = np.vstack((keys[~np.all(keys == fine, axis=1)], children(tuple(fine))))
    with pytest.raises(ValueError):
        LeafMesh(c, keys)  # a level-2 leaf beside a root breaks the 2:1 balance
