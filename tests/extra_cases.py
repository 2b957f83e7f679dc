"""Extra_LU+ known answers and a model whose adaptive-switch level follows from the switch rule
(test fixtures; each answer is derived by hand below, not by running any engine).

Extra_LU+ (Behrmann, Bouyer, Larsen, Pelanek, STTT 8(3) 2006, Def. 6; the reference's
extrapolate_lu_plus, restated in src/metal/engine_metal.mm:80-93), for a canonical zone Z with
entries c_ij bounding x_i - x_j, on every finite off-diagonal entry, clauses in order:
  (1) c_ij > L_i                      -> infinity
  (2) -c_0i > L_i                     -> infinity
  (3) i != 0 and -c_0j > U_j          -> infinity
  (4) i == 0 and -c_0j > U_j          -> (<, -U_j)
then the result is closed. Raw bounds: le(v) = 2v + 1 (x_i - x_j <= v), lt(v) = 2v (< v).
Each case below is built so that one clause decides an entry that no earlier clause touches; a
dropped clause, or an identity extrapolation, leaves that entry (and its consequences after the
closure) different from the expected zone.
"""
from ptctcl_amd.network import INF, Edge, Location, Network, Process, Target, c_le


def le(v):
    return 2 * v + 1


def lt(v):
    return 2 * v


# (name, d, zone row-major, L, U, expected row-major)
EXTRA_CASES = [
    # 2 <= x <= 5, L_x = 3: the upper bound 5 exceeds L_x (clause 1) and goes; x >= 2 stays
    # (-c_01 = 2 is neither above L_0 = 0 in clause 2's sense for row 0, nor above U_x = 10).
    ("clause1", 2, [le(0), le(-2), le(5), le(0)], [0, 3], [0, 10],
     [le(0), le(-2), INF, le(0)]),
    # 7 <= x <= 9, U_x = 5: the lower bound 7 exceeds U_x, so c_01 becomes (<, -5) (clause 4);
    # the upper bound 9 is below L_x = 20 and stays.
    ("clause4", 2, [le(0), le(-7), le(9), le(0)], [0, 20], [0, 5],
     [le(0), lt(-5), le(9), le(0)]),
    # 0 <= x <= 1, 7 <= y <= 9 (so x - y <= -6, y - x <= 9), U_y = 5: y's lower bound 7 exceeds
    # U_y, so c_12 (x - y) goes (clause 3) and c_02 becomes (<, -5) (clause 4). Closing puts
    # c_12 back at c_10 + c_02 = 1 + (<, -5) = (<, -4).
    ("clause3", 3,
     [le(0), le(0), le(-7),
      le(1), le(0), le(-6),
      le(9), le(9), le(0)],
     [0, 20, 20], [0, 20, 5],
     [le(0), le(0), lt(-5),
      le(1), le(0), lt(-4),
      le(9), le(9), le(0)]),
    # 4 <= x <= 5, 3 <= y <= 6, x - y <= 1 (so y - x <= 2), L_x = 3: c_10 = 5 exceeds L_x
    # (clause 1); c_12 = 1 does not, but x's lower bound 4 exceeds L_x, so c_12 goes too
    # (clause 2). Without clause 2, closing would bound x again through c_12 + c_20 = 7.
    ("clause2", 3,
     [le(0), le(-4), le(-3),
      le(5), le(0), le(1),
      le(6), le(2), le(0)],
     [0, 3, 20], [0, 20, 20],
     [le(0), le(-4), le(-3),
      INF, le(0), INF,
      le(6), le(2), le(0)]),
    # 1 <= x <= 2 with L_x = U_x = 10: no clause applies; the zone is returned as it is.
    ("unchanged", 2, [le(0), le(-1), le(2), le(0)], [0, 10], [0, 10],
     [le(0), le(-1), le(2), le(0)]),
]


def switch_model():
    """One process, one location A, clocks x and y. A self-loop A -> A [x == 1, y <= 100] {x := 0},
    and a second self-loop [x < 0, y > 100] that can never fire (x < 0 holds nowhere) but gives y
    the lower-bound constant 100: L and U are 1 for x and 100 for y. The initial zone is
    x = y >= 0; the zone stored at level k is y - x = k, x >= 0 (x is reset when it reaches 1).
    None of these zones is LU-simulated by another: at a point with x = 0 and y = k < 100 a
    simulating point needs the same y (y is below both of its constants), hence x = k - j != 0,
    which no clause allows for x <= 1. So each level adds one live zone to the single discrete state,
    k + 1 after level k. The adaptive rule (src/reachability/engine.cpp:2178-2184) switches
    Extra_LU+ on after the first level at which live zones >= 4 x discrete states: k + 1 >= 4,
    level 3."""
    loc = Location("A")
    step = Edge(0, 0, guard=[(1, 0, le(1)), (0, 1, le(-1)), c_le(2, 100)], resets=[1])
    never = Edge(0, 0, guard=[(1, 0, lt(0)), (0, 2, lt(-100))])
    net = Network(2, [], [], [Process("P", [loc], [step, never], 0)], clock_names=["0", "x", "y"],
                  name="switch").finalize()
    return net, Target(never=True), 3
