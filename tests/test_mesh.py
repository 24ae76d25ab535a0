"""Mesh layer vs the reference's mesh fixtures (tests_mesh/*), CPU only.

The structured generator must reproduce meshing_scripts/create_full_mesh.py
(pinned by the *_gold.dat files the reference's own mesh tests compare
against) and the sample-mesh builder create_sample_mesh.py.
"""
import os
import re

import numpy as np
import pytest

import pressiodemoapps_amd as pda

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref", "mesh")


def _full_cases():
    out = []
    for d in ("mesh_generation_full_1d", "mesh_generation_full_2d", "mesh_generation_full_3d"):
        root = os.path.join(GOLD, d)
        for n in sorted(os.listdir(root)):
            out.append(os.path.join(root, n))
    return out


def _parse(name):
    # mesh_s{stencil}_{nx}x{ny}[x{nz}]_{periodic...}
    m = re.match(r"mesh_s(\d)_(\d+)x(\d+)(?:x(\d+))?_(.*)", name)
    s, nx, ny, nz, per = int(m.group(1)), int(m.group(2)), int(m.group(3)), m.group(4), m.group(5)
    return s, nx, ny, (int(nz) if nz else None), per


@pytest.mark.parametrize("path", _full_cases(), ids=lambda p: os.path.basename(p))
def test_structured_generator_matches_reference_golds(path, tmp_path):
    s, nx, ny, nz, per = _parse(os.path.basename(path))
    if "full_1d" in path:
        n, bounds = (nx,), (0.0, 1.0)
        periodic = (0,) if per == "periodic" else ()
    elif "full_2d" in path:
        n, bounds = (nx, ny), (0.0, 1.0, 0.0, 1.0)
        periodic = tuple(a for a, c in enumerate("xy") if per.startswith("periodic") and c in per.split("_")[1])
    else:
        # tests_mesh/mesh_generation_full_3d/test.cmake:5-8
        n, bounds = (nx, ny, nz), (0.0, 1.0, 0.0, 2.0, 0.0, 3.0)
        periodic = (0, 1, 2) if per == "periodic_xyz" else ()
    mesh = pda.create_structured_mesh(n, bounds, s, periodic)
    out = str(tmp_path / "m")
    mesh.write(out)
    # info.dat: line-by-line identical (the reference test asserts line equality)
    assert open(os.path.join(out, "info.dat")).read() == open(os.path.join(path, "info_gold.dat")).read()
    # connectivity: identical integers
    assert np.array_equal(np.loadtxt(os.path.join(out, "connectivity.dat"), dtype=np.int64),
                          np.loadtxt(os.path.join(path, "connectivity_gold.dat"), dtype=np.int64))
    # coordinates: identical text
    assert open(os.path.join(out, "coordinates.dat")).read() == \
        open(os.path.join(path, "coordinates_gold.dat")).read()
    # and the reader gives back the same mesh
    back = pda.load_cellcentered_uniform_mesh(out)
    assert np.array_equal(back.graph(), mesh.graph())
    assert back.dx() == mesh.dx() and back.dy() == mesh.dy()
    # inverse spacings as the reader computes them, 0 for an absent axis
    for d, inv in ((back.dx(), back.dxInv()), (back.dy(), back.dyInv()), (back.dz(), back.dzInv())):
        assert inv == (1.0 / d if d != 0.0 else 0.0)


def test_sample_mesh_builder_matches_reference(tmp_path):
    d = os.path.join(GOLD, "mesh_generation_sample_2d", "mesh_s3_5x6_nonperiodic_sample")
    full = pda.create_structured_mesh((5, 6), (0.0, 1.0, 0.0, 1.0), 3)
    gids = np.loadtxt(os.path.join(d, "sample_mesh_gids.dat"), dtype=np.int64)
    sm = pda.create_sample_mesh(full, gids)
    out = str(tmp_path / "s")
    sm.write(out)
    for f in ("info.dat", "connectivity.dat", "coordinates.dat", "stencil_mesh_gids.dat"):
        assert open(os.path.join(out, f)).read() == open(os.path.join(d, f)).read(), f
    s, st = sm.sample_gids()
    assert np.array_equal(s, np.sort(gids))
    assert np.array_equal(st, np.loadtxt(os.path.join(d, "stencil_mesh_gids.dat"), dtype=np.int64))


def test_sample_mesh_properties_random():
    """Closed-form invariants of create_sample_mesh.py on a larger random sample:
    stencil = sorted unique(sample U neighbours), graph renumbered consistently."""
    full = pda.create_structured_mesh((17, 13), (0.0, 1.0, 0.0, 1.0), 7, (0,))
    rng = np.random.default_rng(3)
    gids = rng.choice(17 * 13, size=40, replace=False)
    sm = pda.create_sample_mesh(full, np.concatenate([gids, gids[:5]]))  # duplicates collapse
    s, st = sm.sample_gids()
    G = full.graph()
    assert np.array_equal(s, np.unique(gids))
    exp = np.unique(np.concatenate([s, G[s][G[s] >= 0]]))
    assert np.array_equal(st, exp)
    Gs = sm.graph()
    back = np.where(Gs >= 0, st[np.maximum(Gs, 0)], -1)
    assert np.array_equal(back, G[s])
    assert np.array_equal(sm.viewX(), full.viewX()[st])


def test_row_classification_closed_form():
    """inner rows = no missing neighbour within the mesh stencil width
    (impl/mesh_ccu.hpp:385-447)."""
    for s in (3, 5, 7):
        w = (s - 1) // 2
        mesh = pda.create_structured_mesh((11, 9), (0.0, 1.0, 0.0, 1.0), s)
        ni, nn = mesh.row_classes()
        assert ni == (11 - 2 * w) * (9 - 2 * w)
        assert ni + nn == 99


def test_load_errors_are_reported(tmp_path):
    with pytest.raises(pda.PdaError):
        pda.load_cellcentered_uniform_mesh(str(tmp_path / "missing"))
    with pytest.raises(pda.PdaError):
        pda.create_structured_mesh((4, 4), (0, 1, 0, 1), 4)


TAGGING = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref", "mesh",
                       "mesh_boundary_cells_tagging_2d")


@pytest.mark.parametrize("kind", ["fullmesh", "samplemesh"])
@pytest.mark.parametrize("s", [3, 5, 7])
def test_boundary_cell_tagging_golds(kind, s):
    """tests_mesh/mesh_boundary_cells_tagging_2d/main.cc: the cells near the
    boundary (a -1 within the stencil width) and strictly next to it (a
    missing first-order neighbour), as graph(row, 0) ids in row order, on the
    reference's full and sample meshes of stencil 3, 5 and 7."""
    import json
    gold = json.load(open(os.path.join(TAGGING, "gold_gids.json")))
    mesh = pda.load_cellcentered_uniform_mesh(os.path.join(TAGGING, f"{kind}_s{s}"))
    G = mesh.graph()
    near = G[mesh.graphRowsOfCellsNearBd(), 0].tolist()
    strict = G[mesh.graphRowsOfCellsStrictlyOnBd(), 0].tolist()
    if kind == "fullmesh":
        want_near = gold[f"goldCellsGIDsNearBoundaries_fullS{s}"]
        want_strict = gold["goldCellsGIDsStrictlyAtBoundaries_full"]
    else:
        want_near = gold[f"goldCellsGIDsNearBoundaries_sampleS{s}"]
        want_strict = gold[f"goldCellsGIDsStrictlyAtBoundaries_sampleS{s}"]
    assert near == want_near
    assert strict == want_strict
    away = mesh.graphRowsOfCellsAwayFromBd()
    assert len(away) + len(near) == mesh.sampleMeshSize()
    assert mesh.row_classes() == (len(away), len(near))


def test_strictly_on_boundary_is_2d_only():
    m = pda.create_structured_mesh((6, 5, 4), (0, 1, 0, 1, 0, 1), 3)
    with pytest.raises(pda.PdaError):
        m.graphRowsOfCellsStrictlyOnBd()
