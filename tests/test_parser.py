"""Scene loading: the product's host parser (through the C ABI, no GPU) and the
oracle's restatement against the reference's format and its throw sites
(CSGTree.cu:5-152), and against each other bit for bit."""
import numpy as np
import pytest

import csgray as C
import oracle as O
from conftest import scene_names, scene_text
from csgray import scenes as S

# (nodes, primitives, operator depth) of the reference's Test/ scenes.
EXPECTED = {
    "test1.txt": (3, 2, 1), "test2.txt": (3, 2, 1), "test3.txt": (3, 2, 1), "test4.txt": (7, 4, 2),
    "test5.txt": (17, 9, 4), "testSingleSphere.txt": (1, 1, 0), "testSingleCube.txt": (1, 1, 0),
    "testSingleCylinder.txt": (1, 1, 0), "testPipe.txt": (3, 2, 1), "testCubePipe.txt": (5, 3, 2),
    "testWikipedia.txt": (9, 5, 3), "testWikipediaLeft.txt": (3, 2, 1), "testWikipediaRight.txt": (5, 3, 2),
    "testWikipediaMult.txt": (39, 20, 5), "testCubeCutEdges.txt": (25, 13, 5),
    "testCylinderSphere.txt": (11, 6, 3), "testCylinderSpheres.txt": (7, 4, 2),
    "testCylinderSpheres2.txt": (3, 2, 1), "testSphereCutByCubesAndCylinder.txt": (7, 4, 3),
    "testCheese256.txt": (513, 257, 9), "testCheese512.txt": (1025, 513, 10),
}


def both(text):
    return O.OracleScene(text), C.Scene(text, upload=False)


def assert_same_scene(o, s):
    L1, B1, _ = o.nodes()
    P1, _ = o.prims()
    L2, B2, P2 = s.export()
    assert np.array_equal(L1, L2)
    assert np.array_equal(B1.view(np.uint32), B2.view(np.uint32))
    assert np.array_equal(P1.view(np.uint32), P2.view(np.uint32))


def test_all_reference_scenes_present():
    assert sorted(EXPECTED) == scene_names()


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_reference_scene_counts_and_parity(name):
    o, s = both(scene_text(name))
    i = s.info()
    assert (i.nodes, i.primitives, i.op_depth) == EXPECTED[name]
    oi = o.info()
    assert (oi["nodes"], oi["primitives"], oi["op_depth"]) == EXPECTED[name]
    assert_same_scene(o, s)


@pytest.mark.parametrize("text", [
    S.cheese(512, seed=512), S.cheese(64, seed=3), S.stress(255, seed=7),
    S.random_tree(np.random.default_rng(11)), S.random_tree(np.random.default_rng(12)),
])
def test_generated_scene_parity(text):
    o, s = both(text)
    assert_same_scene(o, s)


def test_preorder_links():
    """left(i) = i+1 for every operator; right/parent consistent (CSGTree.cu:27-39)."""
    s = C.Scene(scene_text("test5.txt"), upload=False)
    links, _, _ = s.export()
    for i, (t, p, l, r, par) in enumerate(links):
        if t < 3:
            assert p == -1 and l == i + 1 and r > l
            assert links[l][4] == i and links[r][4] == i
        else:
            assert l == -1 and r == -1 and p >= 0
    assert links[0][4] == -1


BAD = [
    ("", "number of primitives"),
    ("Union\nSphere 0 0 0 FF0000 1", "number of primitives"),
    ("Sphere 0 0 0 FF0000 1\nSphere 0 0 0 FF0000 1", "Cannot parse"),
    ("Sphere 0 0 0 FF000 1", "color"),
    ("Sphere 0 0 0 FF00000 1", "color"),
    ("Sphere 0 0 0 GG0000 1", ""),
    ("Cylinder 0 0 0 FF0000 1 1 361 0 0", "rotX"),
    ("Cylinder 0 0 0 FF0000 1 1 0 -1 0", "rotY"),
    ("Cylinder 0 0 0 FF0000 1 1 0 0 400", "rotZ"),
    ("Cone 0 0 0 FF0000 1", "Unrecognized keyword"),
    ("union\nSphere 0 0 0 FF0000 1\nSphere 0 0 0 FF0000 1", "Unrecognized keyword"),
    ("Sphere 0 0 0 FF0000", "truncated"),
    ("Union\nSphere 0 0 0 FF0000 1\nCylinder 0 0 0 FF0000 1 1 0 0", "truncated"),
    ("Sphere x 0 0 FF0000 1", ""),
    ("Sphere 1e50 0 0 FF0000 1", ""),
    ("Union\nUnion\nSphere 0 0 0 FF0000 1\nSphere 0 0 0 FF0000 1", "number of primitives"),
]


@pytest.mark.parametrize("text,msg", BAD)
def test_rejected_by_both(text, msg):
    with pytest.raises(O.ParseError) as eo:
        O.OracleScene(text)
    with pytest.raises(C.CsgError) as ep:
        C.Scene(text, upload=False)
    assert ep.value.status == C.E_PARSE
    assert msg in str(eo.value) and msg in str(ep.value)


GOOD = [
    "Sphere 1.5abc 0 0 FF0000 1",             # std::stof parses the leading prefix
    "Sphere 0 0 0 ff00aa 1",                  # lower-case hex
    "Union Sphere 0 0 0 FF0000 1 Sphere 1 1 1 00FF00 2",  # no line structure at all
    "\t\tCube\t0 0 0\tFF0000\t2\r\n",           # tabs and CR
    "Cylinder 0 0 0 FF0000 1 1 360 0 0",      # range ends are inclusive
    "Cylinder 0 0 0 FF0000 1 1 0.0 360 12.5",
    "Sphere -0 +2 3e-1 0A0B0C 1e0",
]


@pytest.mark.parametrize("text", GOOD)
def test_accepted_by_both_identically(text):
    o, s = both(text)
    assert_same_scene(o, s)


def test_colour_channels():
    s = C.Scene("Sphere 0 0 0 FF8001 1", upload=False)
    _, _, prims = s.export()
    assert prims[0, 5] == np.float32(1.0)
    assert prims[0, 6] == np.float32(np.float32(128) / np.float32(255))
    assert prims[0, 7] == np.float32(np.float32(1) / np.float32(255))
