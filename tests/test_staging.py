"""Staging lists of coloured plans (the reference plan's ind_map, loc_map,
ind_sizes and ind_offs, pyop2/plan.pyx:112-212), built on the device: for
every block of a plan and every map of the loop, the block's distinct
targets through the map in ascending order and, per element slot and map
column, the position of the slot's target in its block's list, checked
against numpy.unique per block (the reference's own construction) for
block and element colourings, several partition sizes and arities."""

import numpy as np
import pytest

from pyop2_amd.workloads import meshes

pytestmark = pytest.mark.gpu


def _brute(plan, vals, dim):
    out = []
    for b in range(plan.nblocks):
        els = plan.elems[plan.offset[b]:plan.offset[b] + plan.nelems[b]]
        rows = vals[els]
        want = np.unique(rows.ravel())
        out.append((want, rows))
    return out


@pytest.mark.parametrize('colouring,part', [('blocks', 64), ('blocks', 256),
                                            ('elements', 128)])
def test_coloured_plan_staging(hip, colouring, part):
    from pyop2_amd import configuration as cfg, hip as H
    op2 = hip
    mesh = meshes.TriMesh(30, seed=11)
    nodes = op2.Set(mesh.nvert)
    elems = op2.Set(mesh.nelem)
    en = op2.Map(elems, nodes, 3, mesh.elem_node)
    d = op2.Dat(nodes, np.zeros(mesh.nvert), np.float64)
    args = [d(en[0], op2.INC), d(en[1], op2.INC), d(en[2], op2.INC)]
    with cfg.override(hip_colouring=colouring):
        plan = H.get_plan(elems, mesh.nelem, args, part_size=part)
    st = plan.staging(en)
    vals = en._layout_values().reshape(-1, 3)
    assert plan.nblocks > 1
    for b, (want, rows) in enumerate(_brute(plan, vals, 3)):
        o = st['ind_offs'][b]
        assert o == plan.offset[b] * 3
        assert st['ind_sizes'][b] == len(want)
        assert np.array_equal(st['ind_map'][o:o + len(want)], want)
        loc = st['loc_map'][o:o + plan.nelems[b] * 3].reshape(-1, 3)
        assert np.array_equal(want[loc], rows)
