"""CPU: the scene-graph flattening of src/scene.rs:80-181 (instance order = draw order).

SceneGraphIter (scene.rs:146-181) walks the tree depth first in child order.  A Parent pushes
its accumulated model (previous top * model) on `model_stack` and nothing ever pops it
(scene.rs:160-165), so the children that follow a subtree in a parent's list are placed with
the model of the last Parent pushed, not with their own parent's.  These tests pin that rule
and the Child -> Parent promotion of add_child (scene.rs:94-115) in closed form with
translations, whose products are sums of offsets.
"""
import numpy as np

from vtrace_amd.renderer import SceneGraph
from vtrace_amd import scenes


def _t(x, y, z):
    return scenes.model_matrix((x, y, z), 1.0)


def _offsets(g):
    return [(tuple(float(v) for v in m[12:15]), h) for m, h in g]


def test_flat_children_in_order():
    root = SceneGraph(_t(1, 0, 0), None, [SceneGraph.child(_t(0, 2, 0), 7), SceneGraph.child(_t(0, 0, 3), 8)])
    assert _offsets(root) == [((1.0, 2.0, 0.0), 7), ((1.0, 0.0, 3.0), 8)]
    assert root.num_total_children() == 2


def test_siblings_after_a_subtree_inherit_the_last_pushed_model():
    # root(+x1) -> [ A(+y10) -> [c1(+z100)], c2(+x1000) ]
    a = SceneGraph(_t(0, 10, 0), None, [SceneGraph.child(_t(0, 0, 100), 1)])
    root = SceneGraph(_t(1, 0, 0), None, [a, SceneGraph.child(_t(1000, 0, 0), 2)])
    got = _offsets(root)
    assert got[0] == ((1.0, 10.0, 100.0), 1)
    # scoped composition would give (1001, 0, 0); the reference never pops A's model
    assert got[1] == ((1001.0, 10.0, 0.0), 2)


def test_two_nested_subtrees_stack_their_models():
    # root -> [ A(+y1) -> [c1], B(+y10) -> [c2], c3 ]: B is composed with A's model
    a = SceneGraph(_t(0, 1, 0), None, [SceneGraph.child(_t(0, 0, 0), 1)])
    b = SceneGraph(_t(0, 10, 0), None, [SceneGraph.child(_t(0, 0, 0), 2)])
    root = SceneGraph(_t(0, 0, 0), None, [a, b, SceneGraph.child(_t(0, 0, 5), 3)])
    assert _offsets(root) == [((0.0, 1.0, 0.0), 1), ((0.0, 11.0, 0.0), 2), ((0.0, 11.0, 5.0), 3)]
    assert root.num_total_children() == 3


def test_add_child_promotes_a_child_to_a_parent():
    # scene.rs:94-115: the former child stays first with an identity model under the old model
    c = SceneGraph.child(_t(5, 0, 0), 4)
    c.add_child(SceneGraph.child(_t(0, 6, 0), 9))
    assert _offsets(c) == [((5.0, 0.0, 0.0), 4), ((5.0, 6.0, 0.0), 9)]
    assert c.num_total_children() == 2
