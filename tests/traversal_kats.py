"""Hand-derived known answers for the dual-tree traversal.

Two 7-node trees (root 0, children 1 and 2, leaves 3..6; 0-based heap ids,
node k is the reference's node k+1) whose lists were worked out by hand from
the reference recursion `dualtraversefill!` (src/fmm/interaction_lists.jl:23-49):
  - leaf x leaf -> P2P;
  - admissible (max(r1, r2) / R < eta, strict) -> M2L;
  - else a leaf target splits the source, a leaf source splits the target;
  - else the cluster with the LARGER stretched diameter is split, and a tie
    splits the SOURCE (`diameter(t) > diameter(s)` splits the target);
  - children are visited in order (first, second).
Each case also gives particle positions and momenta from which the reference's
median-split tree (clusters.jl:21-49) is exactly the hand-built tree, so the
GPU traversal can be run on the same input.

Case A (stretch 1, eta 0.5), all particles on the x axis:
  node: 0 [0,10]  1 [0,2]  2 [4,10]  3 [0,1]  4 [1,2]  5 [4,5]  6 [5,10]
  half-diagonals (diameter = admissibility radius here):
        0: 5      1: 1     2: 3      3: .5    4: .5    5: .5    6: 2.5
  centres: 5, 1, 7, .5, 1.5, 4.5, 7.5
  (0,0) R=0, tie          -> split source: (0,1), (0,2)
  (0,1) 5/4 >= .5, d0>d1  -> split target: (1,1), (2,1)
    (1,1) tie             -> (1,3), (1,4); each has a leaf source and
                             ratio 1/.5 = 2 -> split target -> P2P (3,3) (4,3) (3,4) (4,4)
    (2,1) 3/6 = .5 exactly, NOT < .5; d2=3 > d1=1 -> split target: (5,1), (6,1)
      (5,1) 1/3.5 < .5 -> M2L      (6,1) 2.5/6.5 < .5 -> M2L
  (0,2) 5/2 >= .5, d0>d2  -> split target: (1,2), (2,2)
    (1,2) 3/6 = .5 exactly; d1=1 < d2=3 -> split source: (1,5), (1,6)
      (1,5) 1/3.5 -> M2L           (1,6) 2.5/6.5 -> M2L
    (2,2) tie             -> (2,5), (2,6); leaf sources, 3/2.5 and 3/.5 -> split target
                          -> P2P (5,5) (6,5) (5,6) (6,6)
  Splitting the smaller cluster instead gives (0,1) -> (0,3), (0,4) and
  M2L (2,3); a tie splitting the target reorders both lists; "<=" in the
  admissibility test makes (2,1) and (1,2) M2L pairs.

Case B (stretch (1, 1, 1.25) from momenta (0, 0, .75): g = sqrt(1.5625); eta .4):
  particles (0,0,0) (1,0,0) (1,0,0) (2,0,0) (3,0,0) (3,0,.6) (3,0,1.2) (3,0,1.8)
  stretched extents: root x 3 vs z 2.25 -> x; node 1 along x, node 2 along z.
  node: 1 x in [0,2] (d1 = 1);  2 z in [0,1.8] (d2 = .9 * 1.25 = 1.125; .9 unstretched)
        3 x [0,1], 4 x [1,2] (r .5);  5 z [0,.6], 6 z [1.2,1.8] (r .375)
  (0,0) -> (0,1), (0,2)
  (0,1) d0 = |(1.5, 0, 1.125)| = 1.875 > d1 -> (1,1), (2,1)
    (1,1) -> P2P (3,3) (4,3) (3,4) (4,4) as in case A
    (2,1) R = |(2, 0, 1.125)| = 2.2947, 1.125/R = .490 >= .4; d2 = 1.125 > d1 = 1
          -> split target: (5,1), (6,1)      [unstretched d2 = .9 < d1: split source]
      (5,1) R = |(2, 0, .375)| = 2.0349, 1/R = .491 -> leaf target: (5,3), (5,4) P2P
      (6,1) R = |(2, 0, 1.875)| = 2.7414, 1/R = .365 < .4 -> M2L
  (0,2) R = 1.5 -> d0 > d2 -> (1,2), (2,2)
    (1,2) .490; d1 < d2 -> split source: (1,5), (1,6)
      (1,5) .491 -> leaf source: (3,5), (4,5) P2P
      (1,6) .365 -> M2L
    (2,2) -> (2,5), (2,6): R = .75, 1.125/.75 -> split target -> P2P (5,5) (6,5) (5,6) (6,6)
"""
import numpy as np

# children[k] = (2k+1, 2k+2) for k = 0, 1, 2; leaves 3..6
CHILDREN = np.array([[1, 2], [3, 4], [5, 6], [-1, -1], [-1, -1], [-1, -1], [-1, -1]],
                    dtype=np.int64)


def _box_x(lo, hi):
    return [lo, 0.0, 0.0, hi, 0.0, 0.0]


CASE_A = dict(
    stretch=[1.0, 1.0, 1.0],
    eta=0.5,
    N0=2,
    bboxes=np.array([_box_x(0, 10), _box_x(0, 2), _box_x(4, 10), _box_x(0, 1), _box_x(1, 2),
                     _box_x(4, 5), _box_x(5, 10)], dtype=float),
    pos=np.array([[x, 0.0, 0.0] for x in (0, 1, 1, 2, 4, 5, 5, 10)], dtype=float),
    mom=np.zeros((8, 3)),
    p2p=[(3, 3), (4, 3), (3, 4), (4, 4), (5, 5), (6, 5), (5, 6), (6, 6)],
    m2l=[(5, 1), (6, 1), (1, 5), (1, 6)],
)

CASE_B = dict(
    stretch=[1.0, 1.0, 1.25],
    eta=0.4,
    N0=2,
    bboxes=np.array([[0, 0, 0, 3, 0, 1.8], [0, 0, 0, 2, 0, 0], [3, 0, 0, 3, 0, 1.8],
                     [0, 0, 0, 1, 0, 0], [1, 0, 0, 2, 0, 0], [3, 0, 0, 3, 0, 0.6],
                     [3, 0, 1.2, 3, 0, 1.8]], dtype=float),
    pos=np.array([[0, 0, 0], [1, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0], [3, 0, 0.6],
                  [3, 0, 1.2], [3, 0, 1.8]], dtype=float),
    mom=np.tile([0.0, 0.0, 0.75], (8, 1)),
    p2p=[(3, 3), (4, 3), (3, 4), (4, 4), (5, 3), (5, 4), (3, 5), (4, 5), (5, 5), (6, 5),
         (5, 6), (6, 6)],
    m2l=[(6, 1), (1, 6)],
)

CASES = {"A": CASE_A, "B": CASE_B}


def hand_tree(case):
    """The hand-built tree in the oracle's tree-dict form (what lists() reads)."""
    return dict(nc=7, children=CHILDREN.copy(), bboxes=case["bboxes"].copy())


def sorted_pairs(pairs):
    a = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    return a[np.lexsort((a[:, 1], a[:, 0]))] if len(a) else a
