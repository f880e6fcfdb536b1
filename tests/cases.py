"""Shared test inputs (seeded problems used by the CPU and GPU tests)."""
from gismo_amd import problems


def zero_block_problems():
    """Problems with a kappa = 0 block: the exact-zero skip of
    gsExprAssembler.h:599 leaves out the entries only the block's elements
    touch.  No function's support lies inside a block, so the matrix stays
    SPD."""
    out = []
    P = problems.box_patch(2, 2, 7)
    problems.perturb_interior(P, 0.15 / 7, 61)
    P.kappa = problems.lognormal_kappa(P.nelem, 0.5, 61)
    problems.zero_kappa_box(P, (2, 3), (3, 4))
    out.append(problems.Problem([P], problems.all_sides(2), [], name="void2d-p2-n7"))
    P = problems.box_patch(3, 2, 5)
    problems.perturb_interior(P, 0.15 / 5, 62)
    P.kappa = problems.lognormal_kappa(P.nelem, 0.5, 62)
    problems.zero_kappa_box(P, (1, 1, 2), (2, 3, 3))
    out.append(problems.Problem([P], problems.all_sides(3), [], name="void3d-p2-n5"))
    P = problems.box_patch(3, 3, 6)
    problems.perturb_interior(P, 0.15 / 6, 63)
    problems.zero_kappa_box(P, (2, 2, 2), (3, 4, 3))
    out.append(problems.Problem([P], problems.all_sides(3), [], name="void3d-p3-n6"))
    # multipatch: a void along the east side of patch 0, the interface with
    # patch 2 (an interface function keeps its entries from the other patch)
    prob = problems.square_grid(2, 2, 4, 2)
    problems.zero_kappa_box(prob.patches[0], (3, 1), (3, 3))
    prob.name = "void-grid2x2-p2-n4"
    out.append(prob)
    return out
