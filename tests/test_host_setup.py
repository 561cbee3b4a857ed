"""Host-side set-up checks that need no GPU: the device-internal numbering.

The reference numbers DOFs by first appearance over elements in mesh order
(utd_util/uts_ls_intf.c:258-308) and every mf_ boundary keeps that numbering;
inside the device the library re-orders rows and element slots for locality.
These tests pin that internal order: it is a permutation, DOFs follow first
appearance over the re-ordered elements, and it shortens the column spread of
the matrix on a randomly numbered mesh.
"""
import numpy as np

import modfem_amd as mf


def _first_appearance(el_nodes, el_order, node_dof, k):
    seen, order = set(), []
    for e in el_order:
        for v in el_nodes[e, :k]:
            d = node_dof[v]
            if d not in seen:
                seen.add(d)
                order.append(d)
    return np.array(order)


def test_device_order_is_first_appearance_over_sorted_elements():
    for make, k in [(mf.Mesh.box_prism, 6), (mf.Mesh.box_tetra, 4)]:
        m = make(5, 4, 3, seed=11, jitter=0.2)
        dof_node, rp, ci = m.setup()
        perm, eperm = m.setup_order()
        a = m.arrays()
        assert np.array_equal(np.sort(perm), np.arange(len(dof_node)))
        assert np.array_equal(np.sort(eperm), np.arange(m.nel))
        node_dof = np.full(a["coor"].shape[0], -1)
        node_dof[dof_node] = np.arange(len(dof_node))
        assert np.array_equal(perm, _first_appearance(a["el_nodes"], eperm, node_dof, k))
        m.free()


def _lines_per_slice(rp, ci, dev_of_row):
    """mean number of distinct 128-byte lines of x (16 doubles) that one
    64-row slice of the matrix gathers, rows and columns in device order"""
    n = len(rp) - 1
    row_of_dev = np.empty_like(dev_of_row)
    row_of_dev[dev_of_row] = np.arange(n)
    counts = []
    for s in range(0, n, 64):
        cols = np.concatenate([dev_of_row[ci[rp[r]:rp[r + 1]]] for r in row_of_dev[s:s + 64]])
        counts.append(len(np.unique(cols // 16)))
    return float(np.mean(counts))


def test_device_order_gathers_fewer_lines():
    # layers of 33^2 nodes: wide enough that the reference order's x window
    # spans many lines per slice
    m = mf.Mesh.box_prism(32, 32, 32, seed=3, jitter=0.2)
    dof_node, rp, ci = m.setup()
    perm, _ = m.setup_order()
    iperm = np.empty_like(perm)
    iperm[perm] = np.arange(len(perm))
    ref = _lines_per_slice(rp, ci, np.arange(len(perm)))
    dev = _lines_per_slice(rp, ci, iperm)
    assert dev < 0.5 * ref, (dev, ref)
    m.free()
