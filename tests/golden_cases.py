"""Set-up of the reference's single-particle known-answer tests.

Mirrors the drivers collocatedTest/staggeredTest of unit_test/tstAPIC.hpp
(:536-842) and unit_test/tstPolyPIC.hpp (:192-500): a 100^3 box with cell
size 0.5 and two halo cells, one particle of mass 0.134 at
(9.31, -8.28, -3.34), and the grid function F(i,j,k,d) of the local index.
Only a small window of the local grid around the particle is materialised.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NODE, CELL, FACE_I, FACE_J, FACE_K, EDGE_I, EDGE_J, EDGE_K = range(8)


def load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


APIC = load("apic_golden.json")
POLY = load("polypic_golden.json")
SETUP = APIC["setup"]
EPS = SETUP["near_eps"]
H = SETUP["cell_size"]
HALO = SETUP["halo_cell_width"]
PM = SETUP["particle_mass"]
PX = np.array(SETUP["particle_x"])
CX = np.array(SETUP["particle_cell_local"])
DT = SETUP["polypic_dt"]

# Window of local indices materialised around the particle.
WIN_LO = CX - 3
WIN_N = np.array([8, 8, 8])


def local_low(box_shift=(0.0, 0.0, 0.0)):
    """Physical position of local node (0,0,0): box low minus the halo."""
    box = np.array(SETUP["global_box"][:3]) + np.array(box_shift)
    return box - HALO * H


def apic_staggered_setup(order, dim):
    """Entity and local low corner of tstAPIC.hpp:staggeredTest.

    Arrays and splines live on Face<dim>; the Location template (Edge<dim>
    for orders 1 and 3, Face<dim> for order 2) only chooses the box shift
    (tstAPIC.hpp:698-716, 745, 776)."""
    shift = np.zeros(3)
    if order in (1, 3):  # Location = Edge<dim>: shift the other dimensions
        for d in range(3):
            if d != dim:
                shift[d] -= 0.25
    else:  # Location = Face<dim>: shift the test dimension
        shift[dim] += 0.25
    return FACE_I + dim, local_low(shift)


def polypic_staggered_setup(dim):
    """tstPolyPIC.hpp:staggeredTest with Location = Edge<dim>: the test
    dimension of the box is shifted by -0.25 (:363-369)."""
    shift = np.zeros(3)
    shift[dim] -= 0.25
    return EDGE_I + dim, local_low(shift)


def grid_function(d):
    """F(i,j,k,d) of tstAPIC.hpp:581-589 on the window, for component d."""
    i = np.arange(WIN_LO[0], WIN_LO[0] + WIN_N[0])[:, None, None]
    j = np.arange(WIN_LO[1], WIN_LO[1] + WIN_N[1])[None, :, None]
    k = np.arange(WIN_LO[2], WIN_LO[2] + WIN_N[2])[None, None, :]
    ic, jc, kc = (i - 2).astype(np.float64), (j - 2).astype(np.float64), (k - 2).astype(np.float64)
    return (0.0000000001 * (d + 1) * ic ** 5 - 0.0000000012 * ((d + 1) + jc * ic) ** 3
            + 0.0000000001 * (ic * jc * kc) ** 2 + float((d + 1) ** 2))


def grid_vector():
    return np.stack([grid_function(d) for d in range(3)], axis=-1)


def stencil_range(order):
    """Local index range checked by the reference for each order."""
    if order == 1:
        return 0, 2
    if order == 2:
        return -1, 2
    return -1, 3


def window_values(arr, order):
    """Values of a window array at the nodes the reference checks, in its
    i-outer/k-inner order (tstAPIC.hpp:70-75 etc.)."""
    a, b = stencil_range(order)
    out = []
    for i in range(CX[0] + a, CX[0] + b):
        for j in range(CX[1] + a, CX[1] + b):
            for k in range(CX[2] + a, CX[2] + b):
                out.append(arr[i - WIN_LO[0], j - WIN_LO[1], k - WIN_LO[2]])
    return np.array(out)
